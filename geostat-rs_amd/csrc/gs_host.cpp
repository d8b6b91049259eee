// Host runtime and C ABI of libgeostat_hip.so.
//
// Mirrors the reference's build-time contract (Kriging::build,
// kriging.rs:262-408; VariogramModel::new, model.rs:554-618; PointSet::new,
// data.rs:57-77; vecchia_plan, vecchia.rs:232-282) on the host, keeps data
// and search index resident in HBM, and drives the gfx950 kernels.  Index
// construction (kd-tree layout, Vecchia ordering and conditioning sets) runs
// here once per handle, as Kriging::new / vecchia_plan do in the reference;
// every per-target / per-point computation runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/geostat_hip.h"
#include "gs_internal.hpp"
#include "gs_maxmin.hpp"
#include "gs_vecchia_mfma.hpp"

using gs::DevModel;

namespace {

thread_local std::string g_err;
// kernel family of the last solve launched on this thread (gs_last_solve_kernel)
thread_local const char *g_solve_kernel = "";

int32_t fail(int32_t code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(e_ == hipErrorOutOfMemory ? GS_ERR_OUT_OF_MEMORY : GS_ERR_DEVICE, \
                        std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

// A fixed device region handed out front to back in 256-byte units
// (the no-allocation entry point's workspace).
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
};

// Device buffer owned by a handle; grows, never shrinks.  With an arena it
// takes its bytes from the arena instead and never allocates or frees.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    Arena *arena = nullptr;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() {
        if (p && !arena) (void)hipFree(p);
    }
    hipError_t ensure(size_t b) {
        if (b <= bytes && p) return hipSuccess;
        if (arena) {
            const size_t need = ((b ? b : 8) + 255) / 256 * 256;
            if (need > arena->cap - arena->used) return hipErrorOutOfMemory;
            p = arena->base + arena->used;
            arena->used += need;
            bytes = b;
            return hipSuccess;
        }
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            bytes = 0;
        }
        hipError_t e = hipMalloc(&p, b ? b : 8);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

inline double to_radians(double deg) { return deg * (M_PI / 180.0); }

// total_cmp key (IEEE totalOrder) for f64.
inline int64_t total_key(double x) {
    int64_t b;
    std::memcpy(&b, &x, 8);
    return b ^ (int64_t)(((uint64_t)(b >> 63)) >> 1);
}

bool aniso_valid(const gs_anisotropy &a) {   // model.rs:376-400
    return a.ratio_z > 0.0 && std::isfinite(a.ratio_z) && a.ratio > 0.0 && std::isfinite(a.ratio) &&
           std::isfinite(a.azimuth_deg) && std::isfinite(a.dip_deg) && std::isfinite(a.rake_deg);
}

// GSLIB setrot matrix (model.rs:407-433), row-major.
void rotation_3d(const gs_anisotropy &a, double r[9]) {
    double alpha = (a.azimuth_deg >= 0.0 && a.azimuth_deg < 270.0) ? to_radians(90.0 - a.azimuth_deg)
                                                                   : to_radians(450.0 - a.azimuth_deg);
    double beta = to_radians(-a.dip_deg);
    double theta = to_radians(a.rake_deg);
    double sina = std::sin(alpha), cosa = std::cos(alpha);
    double sinb = std::sin(beta), cosb = std::cos(beta);
    double sint = std::sin(theta), cost = std::cos(theta);
    double afac1 = 1.0 / a.ratio, afac2 = 1.0 / a.ratio_z;
    r[0] = cosb * cosa;
    r[1] = cosb * sina;
    r[2] = -sinb;
    r[3] = afac1 * (-cost * sina + sint * sinb * cosa);
    r[4] = afac1 * (cost * cosa + sint * sinb * sina);
    r[5] = afac1 * (sint * cosb);
    r[6] = afac2 * (sint * sina + cost * sinb * cosa);
    r[7] = afac2 * (-sint * cosa + cost * sinb * sina);
    r[8] = afac2 * (cost * cosb);
}

// search_point (kriging.rs:16-21) in the exact reference expression order.
void search_xf_host(const gs_anisotropy *a, int D, const double *p, double *out) {
    if (!a) {
        for (int d = 0; d < D; d++) out[d] = p[d];
        return;
    }
    if (D == 3) {
        double r[9];
        rotation_3d(*a, r);
        out[0] = r[0] * p[0] + r[1] * p[1] + r[2] * p[2];
        out[1] = r[3] * p[0] + r[4] * p[1] + r[5] * p[2];
        out[2] = r[6] * p[0] + r[7] * p[1] + r[8] * p[2];
    } else {
        double ang = to_radians(a->azimuth_deg);
        double s = std::sin(ang), c = std::cos(ang);
        out[0] = p[0] * s + p[1] * c;
        out[1] = (p[0] * c - p[1] * s) / a->ratio;
    }
}

gs::SearchXf make_search_xf(const gs_kconfig &cfg, int D) {
    gs::SearchXf X;
    std::memset(&X, 0, sizeof(X));
    if (!cfg.has_anisotropic_search) return X;
    X.on = 1;
    const gs_anisotropy &a = cfg.anisotropic_search;
    double ang = to_radians(a.azimuth_deg);
    X.s = std::sin(ang);
    X.c = std::cos(ang);
    X.ratio = a.ratio;
    if (D == 3) rotation_3d(a, X.r);
    return X;
}

// Lanczos Gamma (bessel.rs:100-123), for the general-nu Matern norm.
double lanczos_gamma(double x) {
    static const double C[9] = {0.9999999999998099, 676.5203681218851, -1259.1392167224028,
                                771.3234287776531, -176.6150291621406, 12.507343278686905,
                                -0.13857109526572012, 9.984369578019572e-6, 1.505632735149312e-7};
    if (x < 0.5) return M_PI / (std::sin(M_PI * x) * lanczos_gamma(1.0 - x));
    double xm1 = x - 1.0, a = C[0], t = xm1 + 7.5;
    for (int i = 1; i < 9; i++) a += C[i] / (xm1 + (double)i);
    return std::sqrt(2.0 * M_PI) * std::pow(t, xm1 + 0.5) * std::exp(-t) * a;
}

// Gauss-Legendre nodes by Newton on P_n (bessel.rs:29-57).
void gauss_legendre(int n, double *x, double *w) {
    int m = (n + 1) / 2;
    for (int i = 0; i < m; i++) {
        double z = std::cos((M_PI * ((double)i + 0.75)) / ((double)n + 0.5));
        for (;;) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 0; j < n; j++) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * (double)j + 1.0) * z * p2 - (double)j * p3) / ((double)j + 1.0);
            }
            double pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z -= p1 / pp;
            if (std::fabs(z - z1) <= 1e-15) {
                w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
                x[i] = -z;
                w[n - 1 - i] = w[i];
                x[n - 1 - i] = z;
                break;
            }
        }
    }
}

double total_sill(const gs_model &m) {   // model.rs:658-660
    double s = -0.0;
    for (int i = 0; i < m.n_structures; i++) s += m.structures[i].sill;
    return m.nugget + s;
}

// VariogramModel::new validation (model.rs:554-618) plus dimension validity
// (model.rs:104-106) and the GPU's supported set.
int32_t validate_model(const gs_model *m, int D, bool allow_power) {
    if (!m) return fail(GS_ERR_INVALID_PARAMETER, "model is NULL");
    if (!(m->nugget >= 0.0) || !std::isfinite(m->nugget))
        return fail(GS_ERR_INVALID_PARAMETER, "nugget must be finite and >= 0");
    if (m->n_structures < 0 || m->n_structures > GS_MAX_STRUCTURES)
        return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: more than 8 nested structures");
    for (int i = 0; i < m->n_structures; i++) {
        const gs_structure &s = m->structures[i];
        if (s.kind < GS_SPHERICAL || s.kind > GS_POWER)
            return fail(GS_ERR_INVALID_PARAMETER, "unknown model kind");
        if (!(s.sill >= 0.0) || !std::isfinite(s.sill))
            return fail(GS_ERR_INVALID_PARAMETER, "partial sill must be finite and >= 0");
        if (!(s.range > 0.0) || !std::isfinite(s.range))
            return fail(GS_ERR_INVALID_PARAMETER, "range must be finite and > 0");
        if (s.kind == GS_MATERN && (!(s.param > 0.0) || s.param > 15.0))
            return fail(GS_ERR_INVALID_PARAMETER, "Matern nu must be finite, > 0 and <= 15");
        if (s.kind == GS_STABLE && (!(s.param > 0.0) || s.param > 2.0))
            return fail(GS_ERR_INVALID_PARAMETER, "Stable alpha must be in (0, 2]");
        if (s.kind == GS_POWER && (!(s.param > 0.0) || s.param >= 2.0))
            return fail(GS_ERR_INVALID_PARAMETER, "Power theta must be in (0, 2)");
        if (s.has_anis && !aniso_valid(s.anis))
            return fail(GS_ERR_INVALID_PARAMETER, "invalid anisotropy");
    }
    if (!(total_sill(*m) > 0.0)) return fail(GS_ERR_INVALID_PARAMETER, "total sill must be positive");
    for (int i = 0; i < m->n_structures; i++) {
        if (m->structures[i].kind == GS_CIRCULAR && D > 2)
            return fail(GS_ERR_INVALID_PARAMETER, "Circular is not a valid covariance in 3 dimensions");
        if (m->structures[i].kind == GS_POWER && !allow_power)
            return fail(GS_ERR_INVALID_PARAMETER,
                        "unsupported on GPU: Power (semivariogram-form) systems");
    }
    return GS_OK;
}

DevModel resolve_model(const gs_model &m, int D, const gs::BesselNodes *d_nodes) {
    DevModel M;
    std::memset(&M, 0, sizeof(M));
    M.nugget = m.nugget;
    M.c0 = total_sill(m);
    M.ns = m.n_structures;
    M.nodes = d_nodes;
    for (int i = 0; i < m.n_structures; i++) {
        const gs_structure &s = m.structures[i];
        gs::DevStructure &o = M.s[i];
        o.kind = s.kind;
        o.param = s.param;
        o.sill = s.sill;
        o.inv_range = 1.0 / s.range;
        o.aniso = s.has_anis ? 1 : 0;
        if (s.has_anis) {
            if (D == 2) {
                double ang = to_radians(s.anis.azimuth_deg);
                double sn = std::sin(ang), cs = std::cos(ang);
                o.m[0] = sn;
                o.m[1] = cs;
                o.m[2] = cs / s.anis.ratio;
                o.m[3] = -sn / s.anis.ratio;
            } else {
                rotation_3d(s.anis, o.m);
            }
        }
        if (s.kind == GS_MATERN) {
            o.mat_norm = std::pow(2.0, 1.0 - s.param) / lanczos_gamma(s.param);
            o.mat_s = std::sqrt(2.0 * s.param);
        }
    }
    return M;
}

// First duplicate location (data.rs:150-163): keys are (c + 0.0).to_bits();
// the first point whose key was seen before, paired with that earlier point.
// Open addressing over point indices (linear probing, load <= 1/2).
bool duplicate_pair(int D, const double *coords, int64_t n, int64_t pair[2]) {
    size_t cap = 16;
    while (cap < (size_t)n * 2) cap <<= 1;
    std::vector<int64_t> tab(cap, -1);
    auto bits = [&](int64_t i, int d) {
        const double c = coords[i * D + d] + 0.0;
        uint64_t b;
        std::memcpy(&b, &c, 8);
        return b;
    };
    for (int64_t i = 0; i < n; i++) {
        uint64_t h = 0x9E3779B97F4A7C15ULL;
        for (int d = 0; d < D; d++) {
            h ^= bits(i, d);
            h *= 0xBF58476D1CE4E5B9ULL;
            h ^= h >> 31;
        }
        for (size_t s = (size_t)h & (cap - 1);; s = (s + 1) & (cap - 1)) {
            const int64_t j = tab[s];
            if (j < 0) {
                tab[s] = i;
                break;
            }
            bool eq = true;
            for (int d = 0; d < D; d++) eq = eq && bits(j, d) == bits(i, d);
            if (eq) {
                pair[0] = j;
                pair[1] = i;
                return true;
            }
        }
    }
    return false;
}

// Median-split kd-tree layout (search.rs:47-54, 102-113).  With distinct
// axis values the layout is unique whatever selection algorithm produced it;
// equal axis values are ordered by index.
void kd_layout(int D, const double *tc, int64_t n, std::vector<uint32_t> &idx) {
    idx.resize((size_t)n);
    for (int64_t i = 0; i < n; i++) idx[(size_t)i] = (uint32_t)i;
    struct Job { int64_t lo, hi; int axis; };
    std::vector<Job> stack;
    stack.push_back({0, n, 0});
    while (!stack.empty()) {
        Job j = stack.back();
        stack.pop_back();
        int64_t len = j.hi - j.lo;
        if (len <= 1) continue;
        int64_t mid = len / 2;
        int axis = j.axis;
        auto less = [&](uint32_t a, uint32_t b) {
            int64_t ka = total_key(tc[(int64_t)a * D + axis]), kb = total_key(tc[(int64_t)b * D + axis]);
            return ka < kb || (ka == kb && a < b);
        };
        std::nth_element(idx.begin() + j.lo, idx.begin() + j.lo + mid, idx.begin() + j.hi, less);
        int na = (axis + 1) % D;
        stack.push_back({j.lo, j.lo + mid, na});
        stack.push_back({j.lo + mid + 1, j.hi, na});
    }
}

int tree_depth(int64_t n) {
    int d = 0;
    while (n > 0) { n >>= 1; d++; }
    return d;
}

int pad_nb(int64_t k) { return k <= 16 ? 16 : (k <= 32 ? 32 : 64); }

// ---- launch schedule table -------------------------------------------------
// Each solve kernel is compiled for two group sizes (lanes per system), and
// list-driven solves run on a persistent grid.  The choice per (kernel
// family, dim, padded size, drift terms, model class, extended systems) comes
// from a table written by the project's autotuner (geostat_hip/tune.py, run
// on an MI355X; gs_tune_table.inc), so it is the same in every process and a
// result never depends on which process measured what: the two group sizes
// reduce in different orders, and their outputs differ in the last bits.
// Shapes missing from the table take group-size variant 0 and 16
// wavefronts per CU.  GS_SOLVE_VARIANT=0|1 pins the variant (the tuner's
// own measurements use it).
struct TuneEntry {
    const char *family;     // "krige" or "vecchia"
    int D, NB, M, lean, ext;
    int variant;            // group-size variant
    int waves_per_cu;       // persistent grid of list-driven solves
};

const TuneEntry kTuneTable[] = {
#include "gs_tune_table.inc"
    {nullptr, 0, 0, 0, 0, 0, 0, 0},
};

struct Schedule {
    int variant = 0;
    int waves_per_cu = 16;
};

Schedule schedule_for(const char *family, int D, int NB, int M, bool lean, bool ext) {
    Schedule sc;
    for (const TuneEntry *e = kTuneTable; e->family; e++) {
        if (std::strcmp(e->family, family) == 0 && e->D == D && e->NB == NB && e->M == M &&
            e->lean == (int)lean && e->ext == (int)ext) {
            sc.variant = e->variant;
            sc.waves_per_cu = e->waves_per_cu;
            break;
        }
    }
    // only the lean, non-extended kernels are compiled for both group sizes
    if (!lean || ext) sc.variant = 0;
    if (const char *env = std::getenv("GS_SOLVE_VARIANT")) sc.variant = std::atoi(env) ? 1 : 0;
    if (const char *env = std::getenv("GS_WAVES_PER_CU")) {
        const int w = std::atoi(env);
        if (w > 0 && w <= 64) sc.waves_per_cu = w;
    }
    return sc;
}

int device_cus(int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    return cus;
}

}  // namespace

struct gs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    DBuf nodes;   // gs::BesselNodes
};

// Per-call device workspace of a predictor: neighbour lists, search order,
// bucket lists, staging for host buffers, global-neighbourhood scratch.
// The handle's own calls use the one inside the handle; the request loop
// gives every slot its own, so requests on one dataset run concurrently.
// A default-priority stream of the workspace's own, with the two events that
// let the kd-tree replay of flagged targets run beside the solve of the
// other targets (created at first use, with the caller's device current).
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t searched = nullptr, replayed = nullptr;
    SideStream() = default;
    SideStream(const SideStream &) = delete;
    SideStream &operator=(const SideStream &) = delete;
    ~SideStream() {
        if (s) (void)hipStreamDestroy(s);
        if (searched) (void)hipEventDestroy(searched);
        if (replayed) (void)hipEventDestroy(replayed);
    }
    hipError_t ensure() {
        if (!s) {
            hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
            if (e != hipSuccess) { s = nullptr; return e; }
        }
        if (!searched) {
            hipError_t e = hipEventCreateWithFlags(&searched, hipEventDisableTiming);
            if (e != hipSuccess) { searched = nullptr; return e; }
        }
        if (!replayed) {
            hipError_t e = hipEventCreateWithFlags(&replayed, hipEventDisableTiming);
            if (e != hipSuccess) { replayed = nullptr; return e; }
        }
        return hipSuccess;
    }
};

struct KWork {
    DBuf nbr, count, lists, lens, h_targets, h_value, h_var, h_lag, h_status;
    DBuf keys, perm, sort_tmp, flags, h_tdrift, blk, g_w, g_b0;
    DBuf mark;          // [T] bytes: 1 for targets flagged for the kd-tree replay
    DBuf sel_tmp;       // selection workspace of the unflagged targets
    DBuf fb;            // [T + 1] FP32-refined fallback list and its length
    // neighbourhoods above 64 points: target list (+ length), lists and
    // counts by list position, device-memory heaps, factorization scratch
    DBuf big_list, big_nbr, big_cnt, big_heap_d, big_heap_ix, big_scratch;
    SideStream side;
    void use_arena(Arena *a) {
        for (DBuf *b : {&nbr, &count, &lists, &lens, &h_targets, &h_value, &h_var, &h_lag,
                        &h_status, &keys, &perm, &sort_tmp, &flags, &h_tdrift,
                        &blk, &g_w, &g_b0, &mark, &sel_tmp, &fb, &big_list, &big_nbr, &big_cnt,
                        &big_heap_d, &big_heap_ix, &big_scratch})
            b->arena = a;
    }
};

struct gs_kriging {
    gs_ctx *ctx = nullptr;
    int D = 2;
    int64_t n = 0;
    gs_kconfig cfg{};
    gs_model model{};
    int M = 1;
    bool lean = true;        // every structure in the fast closed-form class
    int mfma_kind = -1;      // one lean structure: its kind for the matrix-core solve (-1: none)
    bool mfma_aniso = false;
    int mfma64_sig = -1;     // covariance signature of the 64-slot matrix-core solve (-1: none)
    bool has_tree = false;
    int64_t k_heap = 0;
    int per_oct = 0;
    int64_t nb_bound = 0;   // padded system bound (<= 64)
    // Neighbourhoods above 64 points (gs_big.hip): big_req is the list
    // bound of the configuration (0: never above 64); big_only: octant
    // quotas above 64 slots, every target goes through the large path.
    int64_t big_req = 0;
    bool big_only = false;
    int64_t stride = 0;     // neighbour-list capacity per target
    int32_t min_nb = 1;
    int depth = 0;
    double center[3] = {0, 0, 0};
    double scale = 1.0;
    double r2 = std::numeric_limits<double>::infinity();
    gs::SearchXf xf{};
    DBuf coords, values, tcoords, tidx, dmodel;
    // extended systems: measurement-error variances, normalized external
    // covariates (nv of them) and block-kriging workspace
    bool has_err = false;
    // global neighbourhood above 64 points: one shared LU on the device
    bool global = false;
    DBuf g_lu, g_piv, g_flag, g_scale;
    int nv = 0;
    double xmean[3] = {0, 0, 0}, xscale[3] = {1, 1, 1};
    DBuf err, xdrift;
    // cell grid of the search-frame coordinates (tile search)
    DBuf cg_start, cg_pts, cg_idx, cg_rank;   // + Hilbert rank of every cell
    gs::CellGridView cg{};
    // per-call workspace of the calls made through the handle itself
    KWork w;
    std::mutex mu;
    // measurement hook: (search start, search end / solve start, solve end)
    bool timing = false;
    std::vector<std::array<hipEvent_t, 3>> marks;
    std::vector<hipEvent_t> event_pool;
    ~gs_kriging() {
        for (auto &m : marks) event_pool.insert(event_pool.end(), m.begin(), m.end());
        for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
    }
    hipError_t take_event(hipEvent_t *e) {
        if (!event_pool.empty()) {
            *e = event_pool.back();
            event_pool.pop_back();
            return hipSuccess;
        }
        return hipEventCreate(e);
    }
};

struct gs_vplan {
    gs_ctx *ctx = nullptr;
    int D = 2;
    int64_t n = 0, m = 0;
    std::vector<int64_t> order;
    std::vector<int32_t> nb;
    std::vector<int32_t> nb_count;
    // Device storage order: position h of the device arrays holds point
    // perm[h] (points along a space-filling curve), so a block of
    // consecutive positions covers a compact region and the gathers of its
    // conditioning sets share cache lines.  d_order is the identity over
    // positions, d_nb / d_count are the plan rows of the points at those
    // positions, with the conditioning sets as positions, and zc is stored
    // by position.
    std::vector<int64_t> perm;
    int32_t max_count = 0;
    double extent = 0.0;     // length of the bounding-box diagonal (MLE start scale)
    DBuf coords, d_order, d_nb, d_count, zc, models, partial, out, status;
    // FP32-refined batches: per-model fallback flags and the FP64 re-run
    DBuf refine_fail, models_redo, out_redo, status_redo;
    std::mutex mu;
};

int gs_plan_dim(const gs_vplan *p) { return p->D; }
int64_t gs_plan_points(const gs_vplan *p) { return p->n; }
double gs_plan_extent(const gs_vplan *p) { return p->extent; }
int32_t gs_fail(int32_t code, const char *msg) { return fail(code, msg); }

namespace {

// Uniform grid over the search-frame coordinates, about two points per cell
// over the bounding box (axes with ~zero extent get one cell), capped at
// 4n + 64 cells; points stored cell by cell (x fastest).
int32_t build_cell_grid(gs_kriging *K, const std::vector<double> &tc) {
    const int D = K->D;
    const int64_t n = K->n;
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) { mn[d] = INFINITY; mx[d] = -INFINITY; }
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < D; d++) {
            mn[d] = std::fmin(mn[d], tc[(size_t)(i * D + d)]);
            mx[d] = std::fmax(mx[d], tc[(size_t)(i * D + d)]);
        }
    double ext[3] = {0, 0, 0}, max_ext = 0.0;
    for (int d = 0; d < D; d++) { ext[d] = mx[d] - mn[d]; max_ext = std::fmax(max_ext, ext[d]); }
    bool active[3] = {false, false, false};
    int n_active = 0;
    double volume = 1.0;
    for (int d = 0; d < D; d++) {
        active[d] = ext[d] > max_ext * 1e-9 && ext[d] > 0.0;
        if (active[d]) { n_active++; volume *= ext[d]; }
    }
    double cell = 1.0;
    if (n_active > 0) cell = std::pow(volume / std::max(0.5 * (double)n, 1.0), 1.0 / n_active);
    if (!(cell > 0.0) || !std::isfinite(cell)) cell = std::fmax(max_ext, 1.0);
    int64_t nc[3] = {1, 1, 1};
    for (;;) {
        int64_t total = 1;
        for (int d = 0; d < D; d++) {
            nc[d] = active[d] ? std::max<int64_t>(1, (int64_t)std::ceil(ext[d] / cell)) : 1;
            total *= nc[d];
        }
        if (total <= 4 * n + 64 && total < (int64_t)1 << 30) break;
        cell *= 1.25;
    }
    gs::CellGridView &G = K->cg;
    std::memset(&G, 0, sizeof(G));
    for (int d = 0; d < 3; d++) {
        G.origin[d] = d < D ? mn[d] : 0.0;
        G.nc[d] = (int32_t)(d < D ? nc[d] : 1);
    }
    G.cell = cell;
    G.inv_cell = 1.0 / cell;
    const int64_t ncells = nc[0] * nc[1] * nc[2];
    std::vector<int32_t> start((size_t)ncells + 1, 0);
    std::vector<int64_t> cid((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        int64_t id = 0, mul = 1;
        for (int d = 0; d < D; d++) {
            const double f = (tc[(size_t)(i * D + d)] - G.origin[d]) * G.inv_cell;
            int64_t v = f <= 0.0 ? 0 : (f >= (double)(nc[d] - 1) ? nc[d] - 1 : (int64_t)f);
            id += v * mul;
            mul *= nc[d];
        }
        cid[(size_t)i] = id;
        start[(size_t)id + 1]++;
    }
    for (int64_t c = 0; c < ncells; c++) start[(size_t)c + 1] += start[(size_t)c];
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    std::vector<double> pts((size_t)(n * D));
    std::vector<uint32_t> idx((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        const int32_t at = fill[(size_t)cid[(size_t)i]]++;
        idx[(size_t)at] = (uint32_t)i;
        for (int d = 0; d < D; d++) pts[(size_t)at * D + d] = tc[(size_t)(i * D + d)];
    }
    HIP_TRY(K->cg_start.ensure(sizeof(int32_t) * (ncells + 1)));
    HIP_TRY(K->cg_pts.ensure(sizeof(double) * n * D));
    HIP_TRY(K->cg_idx.ensure(sizeof(uint32_t) * n));
    HIP_TRY(hipMemcpy(K->cg_start.p, start.data(), sizeof(int32_t) * (ncells + 1), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(K->cg_pts.p, pts.data(), sizeof(double) * n * D, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(K->cg_idx.p, idx.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    G.start = K->cg_start.as<int32_t>();
    G.pts = K->cg_pts.as<double>();
    G.idx = K->cg_idx.as<uint32_t>();
    HIP_TRY(K->cg_rank.ensure(sizeof(uint32_t) * ncells));
    HIP_TRY(gs::build_cell_ranks(D, G, K->cg_rank.as<uint32_t>(), K->ctx->stream));
    return GS_OK;
}

}  // namespace

namespace {

// ---- fixed scratch of the no-allocation entry point ------------------------
// Every GPU the library runs on holds one copy of this array from module
// load to process end; gs_ctx_create records its device address once per
// device (the array's name is not a device address in host code).
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
std::vector<void *> g_scratch;   // per device

// With `device` current.
int32_t scratch_init(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch.size() <= device) g_scratch.resize((size_t)device + 1, nullptr);
    if (g_scratch[(size_t)device]) return GS_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(GS_ERR_DEVICE, std::string("scratch address: ") + hipGetErrorString(e));
    g_scratch[(size_t)device] = ptr;
    return GS_OK;
}

void *scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < (int)g_scratch.size() ? g_scratch[(size_t)device] : nullptr;
}

}  // namespace

extern "C" {

const char *gs_last_error(void) { return g_err.c_str(); }
const char *gs_last_solve_kernel(void) { return g_solve_kernel; }
int32_t gs_abi_version(void) { return GS_ABI_VERSION; }

int32_t gs_model_validate(const gs_model *model) { return validate_model(model, 2, true); }

int32_t gs_ctx_create(int32_t device, gs_ctx **out) {
    if (!out) return fail(GS_ERR_INVALID_PARAMETER, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(GS_ERR_INVALID_PARAMETER, "no such device");
    HIP_TRY(hipSetDevice(device));
    gs_ctx *c = new gs_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(GS_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    int32_t sst = scratch_init(device);
    if (sst != GS_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return sst;
    }
    gs::BesselNodes nodes;
    gauss_legendre(60, nodes.x60, nodes.w60);
    gauss_legendre(40, nodes.x40, nodes.w40);
    if ((e = c->nodes.ensure(sizeof(nodes))) != hipSuccess ||
        (e = hipMemcpy(c->nodes.p, &nodes, sizeof(nodes), hipMemcpyHostToDevice)) != hipSuccess) {
        gs_ctx_destroy(c);
        return fail(GS_ERR_DEVICE, std::string("bessel nodes: ") + hipGetErrorString(e));
    }
    *out = c;
    return GS_OK;
}

void gs_ctx_destroy(gs_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int32_t gs_ctx_synchronize(gs_ctx *ctx) {
    if (!ctx) return fail(GS_ERR_INVALID_PARAMETER, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GS_OK;
}

}  // extern "C"

namespace {

// PointSet::new checks (data.rs:57-77) and the model's own validation, which
// the reference performs before any predictor constructor runs.
int32_t check_pointset(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                       int64_t n, const gs_model *model, const gs_kconfig *config,
                       gs_kriging **out) {
    if (!out) return fail(GS_ERR_INVALID_PARAMETER, "out is NULL");
    *out = nullptr;
    if (!ctx || !config) return fail(GS_ERR_INVALID_PARAMETER, "ctx/config is NULL");
    if (dim != 2 && dim != 3) return fail(GS_ERR_DIMENSION_MISMATCH, "dim must be 2 or 3");
    if (n <= 0) return fail(GS_ERR_INSUFFICIENT_DATA, "no points provided");
    if (!coords || !values) return fail(GS_ERR_INVALID_PARAMETER, "coords/values is NULL");
    for (int64_t i = 0; i < n * dim; i++)
        if (!std::isfinite(coords[i])) return fail(GS_ERR_INVALID_PARAMETER, "non-finite coordinate");
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite(values[i])) return fail(GS_ERR_INVALID_PARAMETER, "non-finite value");
    return validate_model(model, dim, true);
}

// Largest global neighbourhood: the shared system is (n + M)^2 doubles
// (2 GiB at the limit) and its factorization n + M dependent steps.
constexpr int64_t kGlobalMax = 16384;
// Largest moving neighbourhood solved on the GPU (gs_big.hip); a target
// with more neighbours gets status GS_ERR_INVALID_PARAMETER.
constexpr int64_t kBigCap = 4096;

gs::GlobalParams global_params(gs_kriging *K) {
    gs::GlobalParams g;
    std::memset(&g, 0, sizeof(g));
    g.coords = K->coords.as<double>();
    g.values = K->values.as<double>();
    g.n = K->n;
    g.N = K->n + K->M;
    g.M = K->M;
    g.method = K->cfg.method;
    g.sk_mean = K->cfg.simple_mean;
    for (int d = 0; d < 3; d++) g.center[d] = K->center[d];
    g.scale = K->scale;
    g.err = K->has_err ? K->err.as<double>() : nullptr;
    g.xdrift = K->nv > 0 ? K->xdrift.as<double>() : nullptr;
    g.nv = K->nv;
    for (int k = 0; k < 3; k++) {
        g.xmean[k] = K->xmean[k];
        g.xscale[k] = K->xscale[k];
    }
    g.model = K->dmodel.as<DevModel>();
    g.lu = K->g_lu.as<double>();
    g.piv = K->g_piv.as<int32_t>();
    g.singular = K->g_flag.as<int32_t>();
    return g;
}

// Kriging::build (kriging.rs:262-408).  errors: measurement-error variances
// (n) or NULL; drift: n x nv external covariates or NULL.
int32_t kriging_build(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                      int64_t n, const gs_model *model, const gs_kconfig *config,
                      const double *errors, int32_t nv, const double *drift, gs_kriging **out,
                      int64_t dup_pair[2]) {
    const gs_kconfig &cfg = *config;
    // Kriging::build (kriging.rs:262-339), in the reference's check order.
    if (cfg.method == GS_UNIVERSAL && !(cfg.degree == 1 || cfg.degree == 2))
        return fail(GS_ERR_INVALID_PARAMETER, "universal kriging drift degree must be 1 or 2");
    if (cfg.method < GS_SIMPLE || cfg.method > GS_EXTERNAL_DRIFT)
        return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: kriging method");
    bool power = false;
    for (int i = 0; i < model->n_structures; i++) power |= model->structures[i].kind == GS_POWER;
    if (power && cfg.method == GS_SIMPLE)
        return fail(GS_ERR_INVALID_PARAMETER, "Power models cannot be used with simple kriging");
    if (cfg.has_search_radius && !(cfg.search_radius > 0.0))
        return fail(GS_ERR_INVALID_PARAMETER, "search radius must be positive");
    if (cfg.has_anisotropic_search && !aniso_valid(cfg.anisotropic_search))
        return fail(GS_ERR_INVALID_PARAMETER, "invalid anisotropic_search");
    if (cfg.max_neighbors == 0) return fail(GS_ERR_INVALID_PARAMETER, "max_neighbors must be at least 1");
    int64_t pair[2];
    if (duplicate_pair(dim, coords, n, pair)) {
        if (dup_pair) { dup_pair[0] = pair[0]; dup_pair[1] = pair[1]; }
        return fail(GS_ERR_DUPLICATE_POINTS, "duplicate points " + std::to_string(pair[0]) + ", " +
                                                 std::to_string(pair[1]));
    }
    if (cfg.max_per_octant == 0) return fail(GS_ERR_INVALID_PARAMETER, "max_per_octant must be at least 1");
    if (cfg.min_neighbors != GS_NONE) {
        if (cfg.min_neighbors == 0) return fail(GS_ERR_INVALID_PARAMETER, "min_neighbors must be at least 1");
        if (cfg.min_neighbors > n)
            return fail(GS_ERR_INSUFFICIENT_DATA, "min_neighbors exceeds the number of data points");
        if (cfg.max_neighbors != GS_NONE && cfg.min_neighbors > cfg.max_neighbors)
            return fail(GS_ERR_INVALID_PARAMETER, "min_neighbors exceeds max_neighbors");
    }
    if (power) return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: Power (semivariogram-form) systems");
    if (n > (int64_t)(1u << 29)) return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: more than 2^29 points");

    std::unique_ptr<gs_kriging> K(new gs_kriging());
    K->ctx = ctx;
    K->D = dim;
    K->n = n;
    K->cfg = cfg;
    K->model = *model;
    K->M = cfg.method == GS_SIMPLE ? 0
         : cfg.method == GS_ORDINARY ? 1
         : cfg.method == GS_EXTERNAL_DRIFT ? 1 + nv
         : (cfg.degree == 1 ? 1 + dim : 1 + dim + dim * (dim + 1) / 2);   // kriging.rs:50-59
    K->min_nb = cfg.min_neighbors == GS_NONE ? 1 : (int32_t)cfg.min_neighbors;
    // centre / scale for the drift basis (kriging.rs:340-347)
    double mn[3], mx[3];
    for (int d = 0; d < dim; d++) { mn[d] = INFINITY; mx[d] = -INFINITY; }
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < dim; d++) {
            mn[d] = std::fmin(mn[d], coords[i * dim + d]);
            mx[d] = std::fmax(mx[d], coords[i * dim + d]);
        }
    double spread = 0.0;
    for (int d = 0; d < dim; d++) {
        K->center[d] = 0.5 * (mn[d] + mx[d]);
        spread = std::fmax(spread, mx[d] - mn[d]);
    }
    K->scale = std::fmax(0.5 * spread, std::numeric_limits<double>::min());
    // per-covariate normalization of the external drift (kriging.rs:349-364)
    std::vector<double> xd;
    if (nv > 0) {
        K->nv = nv;
        const double nn = (double)n;
        for (int k = 0; k < nv; k++) {
            double sum = 0.0;
            for (int64_t i = 0; i < n; i++) sum += drift[i * nv + k];
            const double mean = sum / nn;
            double ss = 0.0;
            for (int64_t i = 0; i < n; i++) {
                const double d = drift[i * nv + k] - mean;
                ss += d * d;
            }
            K->xmean[k] = mean;
            K->xscale[k] = std::fmax(std::sqrt(ss / nn), std::numeric_limits<double>::min());
        }
        // basis values at the data, as fill_basis computes them (kriging.rs:431-437)
        xd.resize((size_t)(n * nv));
        for (int64_t i = 0; i < n; i++)
            for (int k = 0; k < nv; k++)
                xd[(size_t)(i * nv + k)] = (drift[i * nv + k] - K->xmean[k]) / K->xscale[k];
    }
    K->has_tree = cfg.max_neighbors != GS_NONE || cfg.has_search_radius || cfg.max_per_octant != GS_NONE;
    // Neighbour-list capacity (stride) and padded system bound (nb_bound).
    // The register-resident solves hold at most 64 neighbours.  When the
    // configuration does not bound the list by 64 (radius-only search,
    // max_neighbors above 64), the fast search keeps 65 slots: a target that
    // fills them goes through the large path (run_big), every other target
    // is exact (its list never saturated).  Octant quotas above 64 slots
    // send every target through the large path.
    if (!K->has_tree) {
        if (n > kGlobalMax)
            return fail(GS_ERR_INVALID_PARAMETER,
                        "unsupported on GPU: global neighbourhood with more than 16384 points");
        K->global = n > 64;
        K->stride = K->global ? 1 : n;
    } else if (cfg.max_per_octant != GS_NONE) {
        const int64_t q = std::min<int64_t>(cfg.max_per_octant, n);
        const int64_t cap = q << dim;
        K->per_oct = (int)q;
        K->stride = cfg.max_neighbors != GS_NONE ? std::min(cap, cfg.max_neighbors) : cap;
        if (cap > 64) {
            if (K->stride > kBigCap)
                return fail(GS_ERR_INVALID_PARAMETER,
                            "unsupported on GPU: octant neighbourhoods above 4096 points");
            K->big_only = true;
            K->big_req = K->stride;
        }
    } else {
        int64_t req = cfg.max_neighbors != GS_NONE ? cfg.max_neighbors : n;
        K->k_heap = std::min<int64_t>(req, 65);
        K->stride = std::min(K->k_heap, n);
        if (std::min(req, n) > 64) K->big_req = std::min(req, n);
    }
    K->nb_bound = std::min<int64_t>(K->stride, 64);
    K->r2 = cfg.has_search_radius ? cfg.search_radius * cfg.search_radius
                                  : std::numeric_limits<double>::infinity();
    K->xf = make_search_xf(cfg, dim);

    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(K->coords.ensure(sizeof(double) * n * dim));
    HIP_TRY(K->values.ensure(sizeof(double) * n));
    HIP_TRY(hipMemcpy(K->coords.p, coords, sizeof(double) * n * dim, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(K->values.p, values, sizeof(double) * n, hipMemcpyHostToDevice));
    if (errors) {
        K->has_err = true;
        HIP_TRY(K->err.ensure(sizeof(double) * n));
        HIP_TRY(hipMemcpy(K->err.p, errors, sizeof(double) * n, hipMemcpyHostToDevice));
    }
    if (nv > 0) {
        HIP_TRY(K->xdrift.ensure(sizeof(double) * n * nv));
        HIP_TRY(hipMemcpy(K->xdrift.p, xd.data(), sizeof(double) * n * nv, hipMemcpyHostToDevice));
    }
    DevModel dm = resolve_model(*model, dim, ctx->nodes.as<gs::BesselNodes>());
    K->lean = gs::model_is_lean(dm);
    K->mfma_kind = gs::vecchia_mfma_kind(&dm, 1, 32, &K->mfma_aniso);
    K->mfma64_sig = gs::krige_mfma64_sig(dm, dim);
    HIP_TRY(K->dmodel.ensure(sizeof(DevModel)));
    HIP_TRY(hipMemcpy(K->dmodel.p, &dm, sizeof(DevModel), hipMemcpyHostToDevice));
    if (K->has_tree) {
        // kd-tree on search-frame coordinates (kriging.rs:367-381)
        std::vector<double> tc((size_t)(n * dim));
        const gs_anisotropy *an = cfg.has_anisotropic_search ? &cfg.anisotropic_search : nullptr;
        for (int64_t i = 0; i < n; i++) search_xf_host(an, dim, coords + i * dim, &tc[(size_t)(i * dim)]);
        std::vector<uint32_t> idx;
        kd_layout(dim, tc.data(), n, idx);
        std::vector<double> tree((size_t)(n * dim));
        for (int64_t p = 0; p < n; p++)
            for (int d = 0; d < dim; d++) tree[(size_t)(p * dim + d)] = tc[(size_t)idx[(size_t)p] * dim + d];
        K->depth = tree_depth(n);
        HIP_TRY(K->tcoords.ensure(sizeof(double) * n * dim));
        HIP_TRY(K->tidx.ensure(sizeof(uint32_t) * n));
        HIP_TRY(hipMemcpy(K->tcoords.p, tree.data(), sizeof(double) * n * dim, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(K->tidx.p, idx.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
        int32_t st2 = build_cell_grid(K.get(), tc);
        if (st2 != GS_OK) return st2;
    }
    if (K->global && n >= K->M) {
        // factor the target-independent system once (kriging.rs:396-406)
        const int64_t N = n + K->M;
        HIP_TRY(K->g_lu.ensure(sizeof(double) * N * N));
        HIP_TRY(K->g_piv.ensure(sizeof(int32_t) * N));
        HIP_TRY(K->g_flag.ensure(sizeof(int32_t)));
        HIP_TRY(K->g_scale.ensure(sizeof(double)));
        hipStream_t s = ctx->stream;
        HIP_TRY(gs::launch_zero_i32(K->g_flag.as<int32_t>(), 1, s));
        gs::GlobalParams gp = global_params(K.get());
        HIP_TRY(gs::launch_global_factor(dim, K->lean, gp, K->g_scale.as<double>(), s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    *out = K.release();
    return GS_OK;
}

}  // namespace

extern "C" {

int32_t gs_kriging_new(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                       int64_t n, const gs_model *model, const gs_kconfig *config, gs_kriging **out,
                       int64_t dup_pair[2]) {
    int32_t st = check_pointset(ctx, dim, coords, values, n, model, config, out);
    if (st != GS_OK) return st;
    if (config->method == GS_EXTERNAL_DRIFT)
        return fail(GS_ERR_INVALID_PARAMETER,
                    "unsupported on GPU: external drift needs gs_kriging_new_with_external_drift");
    return kriging_build(ctx, dim, coords, values, n, model, config, nullptr, 0, nullptr, out,
                         dup_pair);
}

int32_t gs_kriging_new_with_measurement_error(gs_ctx *ctx, int32_t dim, const double *coords,
                                              const double *values, int64_t n,
                                              const gs_model *model, const gs_kconfig *config,
                                              const double *errors, gs_kriging **out,
                                              int64_t dup_pair[2]) {
    int32_t st = check_pointset(ctx, dim, coords, values, n, model, config, out);
    if (st != GS_OK) return st;
    // with_measurement_error (kriging.rs:181-218), in the reference's order
    if (config->method == GS_EXTERNAL_DRIFT)
        return fail(GS_ERR_INVALID_PARAMETER,
                    "measurement error with external drift is not supported yet");
    for (int i = 0; i < model->n_structures; i++)
        if (model->structures[i].kind == GS_POWER)
            return fail(GS_ERR_INVALID_PARAMETER,
                        "measurement error with Power (IRF-0) models is not supported yet");
    if (!errors) return fail(GS_ERR_DIMENSION_MISMATCH, "no error variances for the data points");
    for (int64_t i = 0; i < n; i++)
        if (!(std::isfinite(errors[i]) && errors[i] >= 0.0))
            return fail(GS_ERR_INVALID_PARAMETER,
                        "measurement-error variances must be finite and non-negative");
    return kriging_build(ctx, dim, coords, values, n, model, config, errors, 0, nullptr, out,
                         dup_pair);
}

int32_t gs_kriging_new_with_external_drift(gs_ctx *ctx, int32_t dim, const double *coords,
                                           const double *values, int64_t n,
                                           const gs_model *model, const gs_kconfig *config,
                                           int32_t n_vars, const double *drift_data,
                                           gs_kriging **out, int64_t dup_pair[2]) {
    int32_t st = check_pointset(ctx, dim, coords, values, n, model, config, out);
    if (st != GS_OK) return st;
    // with_external_drift (kriging.rs:223-260), in the reference's order
    if (config->method != GS_EXTERNAL_DRIFT)
        return fail(GS_ERR_INVALID_PARAMETER,
                    "with_external_drift requires KrigingMethod::ExternalDrift");
    if (n_vars <= 0)
        return fail(GS_ERR_INVALID_PARAMETER, "external drift requires at least one covariate");
    if (!drift_data) return fail(GS_ERR_DIMENSION_MISMATCH, "no drift rows for the data points");
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < n_vars; k++)
            if (!std::isfinite(drift_data[i * n_vars + k]))
                return fail(GS_ERR_INVALID_PARAMETER,
                            "non-finite drift value at data point " + std::to_string(i));
    if (n_vars > 3)
        return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: more than 3 external covariates");
    return kriging_build(ctx, dim, coords, values, n, model, config, nullptr, n_vars, drift_data,
                         out, dup_pair);
}

void gs_kriging_free(gs_kriging *k) {
    if (!k) return;
    (void)hipSetDevice(k->ctx->device);
    delete k;
}

int64_t gs_kriging_max_neighbors(const gs_kriging *k) {
    if (!k) return -1;
    if (k->global) return k->n;
    if (k->big_req > 0) return std::min<int64_t>(k->big_req, kBigCap);
    return k->stride;
}

int32_t gs_kriging_set_timing(gs_kriging *k, int32_t enable) {
    if (!k) return fail(GS_ERR_INVALID_PARAMETER, "handle is NULL");
    std::lock_guard<std::mutex> lk(k->mu);
    k->timing = enable != 0;
    return GS_OK;
}

int32_t gs_kriging_timing_summary(gs_kriging *k, int64_t *calls, double *ms_search,
                                  double *ms_solve) {
    if (!k) return fail(GS_ERR_INVALID_PARAMETER, "handle is NULL");
    std::lock_guard<std::mutex> lk(k->mu);
    HIP_TRY(hipSetDevice(k->ctx->device));
    double a = 0.0, b = 0.0;
    for (auto &m : k->marks) {
        HIP_TRY(hipEventSynchronize(m[2]));
        float x = 0.f, y = 0.f;
        HIP_TRY(hipEventElapsedTime(&x, m[0], m[1]));
        HIP_TRY(hipEventElapsedTime(&y, m[1], m[2]));
        a += x;
        b += y;
    }
    if (calls) *calls = (int64_t)k->marks.size();
    if (ms_search) *ms_search = a;
    if (ms_solve) *ms_solve = b;
    for (auto &m : k->marks) k->event_pool.insert(k->event_pool.end(), m.begin(), m.end());
    k->marks.clear();
    return GS_OK;
}

}  // extern "C"

namespace {

int32_t stride_of(const gs_kriging *K) { return (int32_t)std::max<int64_t>(K->stride, 1); }
int64_t ncells_of(const gs_kriging *K) {
    return (int64_t)K->cg.nc[0] * K->cg.nc[1] * K->cg.nc[2];
}

// Neighbour lists for T device targets into the handle's nbr/count buffers.
// side: the kd-tree replay of the flagged targets goes onto the workspace's
// side stream after the tile search (W.side.replayed marks its end) and the
// flagged targets are marked in W.mark; otherwise it follows on s.
int32_t run_neighbors(gs_kriging *K, KWork &W, const double *d_targets, int64_t T, hipStream_t s,
                      bool side = false) {
    const int32_t stride = stride_of(K);
    HIP_TRY(W.nbr.ensure(sizeof(int32_t) * std::max<int64_t>(T, 1) * stride));
    HIP_TRY(W.count.ensure(sizeof(int32_t) * std::max<int64_t>(T, 1)));
    if (!K->has_tree) {
        HIP_TRY(gs::launch_iota_nbr(W.nbr.as<int32_t>(), W.count.as<int32_t>(), T, (int32_t)K->n,
                                    stride, s));
        return GS_OK;
    }
    // 1) Hilbert order of the targets' search-frame cells
    const size_t tmp_bytes = gs::tile_sort_workspace(T, ncells_of(K));
    HIP_TRY(W.keys.ensure(sizeof(uint32_t) * T));
    HIP_TRY(W.perm.ensure(sizeof(int32_t) * T));
    HIP_TRY(W.sort_tmp.ensure(tmp_bytes));
    HIP_TRY(W.flags.ensure(sizeof(int32_t) * (T + 4)));
    HIP_TRY(gs::launch_tile_order(K->D, d_targets, T, K->xf, K->cg, K->cg_rank.as<uint32_t>(),
                                  W.keys.as<uint32_t>(), W.perm.as<int32_t>(), W.sort_tmp.p,
                                  tmp_bytes, s));
    // 2) tile search: canonical lists; boundary ties and scattered
    //    wavefronts flagged
    int32_t *flag_len = W.flags.as<int32_t>() + T;
    HIP_TRY(gs::launch_zero_i32(flag_len, 1, s));
    gs::TileParams tp;
    std::memset(&tp, 0, sizeof(tp));
    tp.G = K->cg;
    tp.targets = d_targets;
    tp.T = T;
    tp.perm = W.perm.as<int32_t>();
    tp.k = (int32_t)K->k_heap;
    tp.per_octant = K->per_oct;
    tp.max_nb = (int32_t)K->stride;
    tp.stride = stride;
    tp.r2 = K->r2;
    tp.xf = K->xf;
    tp.out_idx = W.nbr.as<int32_t>();
    tp.out_count = W.count.as<int32_t>();
    tp.flag_list = W.flags.as<int32_t>();
    tp.flag_len = flag_len;
    if (side) {
        HIP_TRY(W.mark.ensure((size_t)((T + 3) / 4) * 4));
        HIP_TRY(gs::launch_zero_i32(W.mark.as<int32_t>(), (T + 3) / 4, s));
        tp.flag_mark = W.mark.as<uint8_t>();
    }
    HIP_TRY(gs::launch_tile_knn(K->D, tp, s));
    // 3) exact kd-tree replay (the reference's traversal order) of the
    //    flagged targets
    gs::KnnParams p;
    std::memset(&p, 0, sizeof(p));
    p.list = W.flags.as<int32_t>();
    p.list_len = flag_len;
    p.tcoords = K->tcoords.as<double>();
    p.tidx = K->tidx.as<uint32_t>();
    p.n = K->n;
    p.targets = d_targets;
    p.T = T;
    p.k = (int32_t)K->k_heap;
    p.per_octant = K->per_oct;
    p.max_nb = (int32_t)K->stride;
    p.depth = K->depth;
    p.r2 = K->r2;
    p.xf = K->xf;
    p.out_idx = W.nbr.as<int32_t>();
    p.out_count = W.count.as<int32_t>();
    p.stride = stride;
    if (side) {
        HIP_TRY(W.side.ensure());
        HIP_TRY(hipEventRecord(W.side.searched, s));
        HIP_TRY(hipStreamWaitEvent(W.side.s, W.side.searched, 0));
        HIP_TRY(gs::launch_knn(K->D, p, W.side.s));
        HIP_TRY(hipEventRecord(W.side.replayed, W.side.s));
        return GS_OK;
    }
    HIP_TRY(gs::launch_knn(K->D, p, s));
    return GS_OK;
}

// Per-call inputs of the extended systems: target covariates (external
// drift) and the block discretization with its device-side within-block
// covariance.
struct Extra {
    const double *tdrift = nullptr;
    const double *offsets = nullptr;
    int32_t n_off = 0;
    const double *cbb = nullptr;
};

int32_t run_solves(gs_kriging *K, KWork &W, const double *d_targets, int64_t T,
                   int32_t precision, double *d_value, double *d_var, double *d_lag,
                   int32_t *d_status, hipStream_t s, const Extra &x, bool side = false);

// Global neighbourhood above 64 points: per-target substitutions with the
// shared factors, in chunks whose scratch (b0 and the solution, N doubles
// each per target) stays within 1 GiB.
int32_t run_global(gs_kriging *K, KWork &W, const double *d_targets, int64_t T, double *d_value,
                   double *d_var, double *d_lag, int32_t *d_status, hipStream_t s,
                   const Extra &x) {
    const int64_t N = K->n + K->M;
    if (K->n < K->M) {   // n < m drift terms: InsufficientData at every target (kriging.rs:579-584)
        const double nan = std::numeric_limits<double>::quiet_NaN();
        HIP_TRY(gs::launch_fill_f64(d_value, T, nan, s));
        HIP_TRY(gs::launch_fill_f64(d_var, T, nan, s));
        if (d_lag) HIP_TRY(gs::launch_fill_f64(d_lag, T, nan, s));
        if (d_status) HIP_TRY(gs::launch_fill_i32(d_status, T, GS_ERR_INSUFFICIENT_DATA, s));
        return GS_OK;
    }
    int64_t tc = ((int64_t)1 << 30) / (2 * N * (int64_t)sizeof(double));
    tc = std::max<int64_t>(64, tc / 64 * 64);
    tc = std::min<int64_t>(tc, (T + 63) / 64 * 64);
    HIP_TRY(W.g_w.ensure(sizeof(double) * N * tc));
    HIP_TRY(W.g_b0.ensure(sizeof(double) * N * tc));
    gs::GlobalParams g = global_params(K);
    g.targets = d_targets;
    g.T = T;
    g.stride = tc;
    g.w = W.g_w.as<double>();
    g.b0 = W.g_b0.as<double>();
    g.value = d_value;
    g.variance = d_var;
    g.lagrange = d_lag;
    g.status = d_status;
    g.tdrift = x.tdrift;
    g.offsets = x.offsets;
    g.n_off = x.n_off;
    g.cbb = x.cbb;
    for (int64_t t0 = 0; t0 < T; t0 += tc) {
        g.t0 = t0;
        g.Tc = std::min<int64_t>(tc, T - t0);
        HIP_TRY(gs::launch_global_solve(K->D, K->lean, g, s));
    }
    return GS_OK;
}

int32_t run_big(gs_kriging *K, KWork &W, const double *d_targets, int64_t T, double *d_value,
                double *d_var, double *d_lag, int32_t *d_status, hipStream_t s, const Extra &x);

int32_t run_predict(gs_kriging *K, KWork &W, const double *d_targets, int64_t T,
                    int32_t precision, double *d_value, double *d_var, double *d_lag,
                    int32_t *d_status, hipStream_t s, const Extra &x = Extra(),
                    bool timed = true, bool allow_side = true) {
    if (precision != GS_PREC_FP64 && precision != GS_PREC_FP32_REFINED)
        return fail(GS_ERR_INVALID_PARAMETER, "unknown precision");
    if (T == 0) return GS_OK;
    if (K->global)   // one shared FP64 factorization serves either precision
        return run_global(K, W, d_targets, T, d_value, d_var, d_lag, d_status, s, x);
    std::array<hipEvent_t, 3> ev{};
    if (timed && K->timing) {
        for (auto &e : ev) HIP_TRY(K->take_event(&e));
        HIP_TRY(hipEventRecord(ev[0], s));
    }
    // The kd-tree replay of flagged targets overlaps the solve of the
    // others.  Small calls keep it in line: the extra launches and events
    // would cost more than the overlap saves.
    const bool side = allow_side && K->has_tree && T >= 65536;
    int32_t st = GS_OK;
    if (!K->big_only) st = run_neighbors(K, W, d_targets, T, s, side);
    if (st != GS_OK) return st;
    if (timed && K->timing) HIP_TRY(hipEventRecord(ev[1], s));
    if (!K->big_only)
        st = run_solves(K, W, d_targets, T, precision, d_value, d_var, d_lag, d_status, s, x, side);
    if (st != GS_OK) return st;
    // neighbourhoods above 64 points, in FP64 whatever the precision
    if (K->big_req > 0) st = run_big(K, W, d_targets, T, d_value, d_var, d_lag, d_status, s, x);
    if (st != GS_OK) return st;
    if (timed && K->timing) {
        HIP_TRY(hipEventRecord(ev[2], s));
        K->marks.push_back(ev);
    }
    return GS_OK;
}

template <class Solve>
int32_t solve_all(gs_kriging *K, KWork &W, gs::SolveParams p, int64_t T, hipStream_t s, bool side,
                  Solve &&solve);

// Smallest FP32 pivot trusted by the refined kriging path, relative to the
// system size times the matrix scale: the FP32 solve error is about
// cond * 6e-8, the refinement takes it to its square, and cond is at most
// about N * scale / (smallest pivot), so 6e-4 keeps cond * 6e-8 <= 1e-4 and
// the refined error near 1e-8.
constexpr double kFallbackRatioKrige = 6e-4;
// The Vecchia conditionals only enter the likelihood through sums of logs
// and squares over many points: cond * 6e-8 <= 1e-3 (refined per-point
// error near 1e-6, averaged over the points).
constexpr double kFallbackRatioVecchia = 6e-5;

// Kernel parameters of a predict call (data, model, method, outputs and the
// extended inputs); neighbour lists from the workspace.
gs::SolveParams solve_params(gs_kriging *K, KWork &W, const double *d_targets, int64_t T,
                             double *d_value, double *d_var, double *d_lag, int32_t *d_status,
                             const Extra &x) {
    gs::SolveParams p;
    std::memset(&p, 0, sizeof(p));
    p.coords = K->coords.as<double>();
    p.values = K->values.as<double>();
    p.targets = d_targets;
    p.nbr = W.nbr.as<int32_t>();
    p.count = W.count.as<int32_t>();
    p.T = T;
    p.stride = stride_of(K);
    p.min_nb = K->min_nb;
    p.method = K->cfg.method;
    p.degree = K->cfg.degree;
    p.sk_mean = K->cfg.simple_mean;
    for (int d = 0; d < 3; d++) p.center[d] = K->center[d];
    p.scale = K->scale;
    p.model = K->dmodel.as<DevModel>();
    p.value = d_value;
    p.variance = d_var;
    p.lagrange = d_lag;
    p.status = d_status;
    p.err = K->has_err ? K->err.as<double>() : nullptr;
    p.xdrift = K->nv > 0 ? K->xdrift.as<double>() : nullptr;
    p.tdrift = x.tdrift;
    for (int k = 0; k < K->nv; k++) {
        p.xmean[k] = K->xmean[k];
        p.xscale[k] = K->xscale[k];
    }
    p.offsets = x.offsets;
    p.n_off = x.n_off;
    p.cbb = x.cbb;
    return p;
}

int32_t run_solves(gs_kriging *K, KWork &W, const double *d_targets, int64_t T,
                   int32_t precision, double *d_value, double *d_var, double *d_lag,
                   int32_t *d_status, hipStream_t s, const Extra &x, bool side) {
    gs::SolveParams p = solve_params(K, W, d_targets, T, d_value, d_var, d_lag, d_status, x);
    const bool ext = p.err || p.xdrift || p.offsets;
    // FP32-refined solves cover lean models with simple or ordinary
    // kriging; every other configuration is solved in FP64, which meets the
    // refined path's tolerance by itself.
    const bool r32 = precision == GS_PREC_FP32_REFINED && K->lean && !ext && K->M <= 1;
    if (r32) {
        HIP_TRY(W.fb.ensure(sizeof(int32_t) * (T + 1)));
        p.fb_list = W.fb.as<int32_t>();
        p.fb_len = W.fb.as<int32_t>() + T;
        p.fb_ratio = kFallbackRatioKrige;
        HIP_TRY(gs::launch_zero_i32(p.fb_len, 1, s));
    }
    const int cus = device_cus(K->ctx->device);
    // persistent = false: one group per system; otherwise a persistent grid
    // of the table's wavefronts per CU, grid-strided over the list
    // 17-32-neighbour simple and ordinary kriging with one lean structure:
    // the factorization on the matrix cores (GS_KRIGE_MFMA=0 keeps the
    // vector-unit kernel, for comparison)
    const char *mf_env = std::getenv("GS_KRIGE_MFMA");
    const bool mfma = K->mfma_kind >= 0 && K->M <= 1 && !ext && !(mf_env && mf_env[0] == '0');
    // 33-64 neighbours: one system per wavefront on the matrix cores, for
    // the covariance signatures the library holds a kernel for
    const bool mfma64 = K->mfma64_sig >= 0 && !ext && !(mf_env && mf_env[0] == '0') &&
                        gs::has_krige_mfma64(K->D, K->mfma64_sig, K->M);
    auto solve64 = [&](int NB, const gs::SolveParams &pp, bool persistent) -> int32_t {
        const Schedule sc = schedule_for("krige", K->D, NB, K->M, K->lean, ext);
        if (mfma64 && NB == 64) {
            // two wavefronts per SIMD are resident; each strides over the systems
            const int64_t groups = std::min<int64_t>(pp.T, (int64_t)cus * 8);
            HIP_TRY(gs::launch_krige_mfma64(K->D, K->mfma64_sig, K->M, pp, groups, s));
            g_solve_kernel = "krige_mfma64_kernel";
            return GS_OK;
        }
        if (mfma && NB == 32) {
            int64_t groups = pp.T;
            if (persistent) groups = std::min<int64_t>(pp.T, (int64_t)cus * sc.waves_per_cu * 4);
            HIP_TRY(gs::launch_krige_mfma(K->D, K->mfma_aniso, K->mfma_kind, K->M, pp, groups, s));
            g_solve_kernel = "krige_mfma_kernel";
            return GS_OK;
        }
        const int gpw = 64 / gs::group_size(NB, sc.variant);
        int64_t groups = pp.T;
        if (persistent) groups = std::min<int64_t>(pp.T, (int64_t)cus * sc.waves_per_cu * gpw);
        hipError_t e = gs::launch_krige_fp64(K->D, NB, K->M, sc.variant, K->lean, ext, pp, groups, s);
        if (e == hipErrorInvalidValue)
            return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: no kernel for this system size");
        HIP_TRY(e);
        g_solve_kernel = "krige_fp64_kernel";
        return GS_OK;
    };
    auto solve = [&](int NB, const gs::SolveParams &pp, bool persistent) -> int32_t {
        if (!r32) return solve64(NB, pp, persistent);
        const Schedule sc = schedule_for("krige32r", K->D, NB, K->M, true, false);
        const int gpw = 64 / gs::fp32r_group(NB);
        int64_t groups = pp.T;
        if (persistent) groups = std::min<int64_t>(pp.T, (int64_t)cus * sc.waves_per_cu * gpw);
        hipError_t e = gs::launch_krige_fp32r(K->D, NB, K->M, pp, groups, s);
        if (e == hipErrorInvalidValue)
            return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: no kernel for this system size");
        HIP_TRY(e);
        g_solve_kernel = "krige_fp32r_kernel";
        return GS_OK;
    };
    int32_t st = solve_all(K, W, p, T, s, side, solve);
    if (st != GS_OK || !r32) return st;
    // systems the FP32 factorization did not trust, in FP64 (written in
    // place; the list order does not affect any result)
    gs::SolveParams pf = p;
    pf.list = p.fb_list;
    pf.list_len = p.fb_len;
    const char *main_kernel = g_solve_kernel;
    st = solve64(pad_nb(K->nb_bound), pf, true);
    g_solve_kernel = main_kernel;
    return st;
}

// The launch sequence of run_solves for one precision: size buckets for
// uneven octant neighbourhoods, Hilbert order otherwise, and the flagged
// targets after the kd-tree replay when it runs on the side stream.
template <class Solve>
int32_t solve_all(gs_kriging *K, KWork &W, gs::SolveParams p, int64_t T, hipStream_t s, bool side,
                  Solve &&solve) {
    if (K->per_oct > 0 && K->nb_bound > 16) {
        // Uneven sizes: bucket by padded system size, one launch per bucket.
        HIP_TRY(W.lists.ensure(sizeof(int32_t) * 3 * T));
        HIP_TRY(W.lens.ensure(sizeof(int32_t) * 8));
        auto bucket_and_solve = [&](const int32_t *perm, const int32_t *n_dev,
                                    const uint8_t *skip) -> int32_t {
            HIP_TRY(gs::launch_zero_i32(W.lens.as<int32_t>(), 4, s));
            HIP_TRY(gs::launch_bucket(W.count.as<int32_t>(), perm, T, W.lists.as<int32_t>(),
                                      W.lens.as<int32_t>(), T, s, n_dev, skip));
            const int nbs[3] = {16, 32, 64};
            for (int b = 0; b < 3; b++) {
                if (nbs[b] > pad_nb(K->nb_bound)) break;
                gs::SolveParams pb = p;
                pb.list = W.lists.as<int32_t>() + (int64_t)b * T;
                pb.list_len = W.lens.as<int32_t>() + b;
                int32_t st = solve(nbs[b], pb, true);
                if (st != GS_OK) return st;
            }
            return GS_OK;
        };
        if (side) {
            // the unflagged targets while the replay runs, then the flagged
            // ones (same lists, reused in stream order)
            int32_t st = bucket_and_solve(W.perm.as<int32_t>(), nullptr, W.mark.as<uint8_t>());
            if (st != GS_OK) return st;
            HIP_TRY(hipStreamWaitEvent(s, W.side.replayed, 0));
            return bucket_and_solve(W.flags.as<int32_t>(), W.flags.as<int32_t>() + T, nullptr);
        }
        return bucket_and_solve(K->has_tree ? W.perm.as<int32_t>() : nullptr, nullptr, nullptr);
    }
    if (K->has_tree) {
        // process targets in the search's Hilbert order: neighbouring systems
        // share data points, so the coordinate gathers hit cache
        HIP_TRY(W.lens.ensure(sizeof(int32_t) * 8));
        HIP_TRY(gs::launch_fill_i32(W.lens.as<int32_t>() + 3, 1, (int32_t)T, s));
        p.list = W.perm.as<int32_t>();
        p.list_len = W.lens.as<int32_t>() + 3;
        if (side) {
            // the unflagged targets while the replay runs, then the flagged
            // ones once their lists are in
            HIP_TRY(W.lists.ensure(sizeof(int32_t) * T));
            const size_t sel_bytes = gs::select_workspace(T);
            HIP_TRY(W.sel_tmp.ensure(sel_bytes));
            HIP_TRY(gs::launch_select_unmarked(W.perm.as<int32_t>(), W.mark.as<uint8_t>(), T,
                                               W.lists.as<int32_t>(), W.lens.as<int32_t>() + 5,
                                               W.sel_tmp.p, sel_bytes, s));
            p.list = W.lists.as<int32_t>();
            p.list_len = W.lens.as<int32_t>() + 5;
            int32_t st = solve(pad_nb(K->nb_bound), p, false);
            if (st != GS_OK) return st;
            HIP_TRY(hipStreamWaitEvent(s, W.side.replayed, 0));
            gs::SolveParams pf = p;
            pf.list = W.flags.as<int32_t>();
            pf.list_len = W.flags.as<int32_t>() + T;
            return solve(pad_nb(K->nb_bound), pf, true);
        }
    }
    return solve(pad_nb(K->nb_bound), p, false);
}

// ---- neighbourhoods above 64 points ----------------------------------------
// The targets that need the large path: every target of a big_only handle,
// otherwise those whose fast search filled its 65 slots.  *F on the host.
int32_t big_targets(gs_kriging *K, KWork &W, int64_t T, hipStream_t s, int64_t *F) {
    HIP_TRY(W.big_list.ensure(sizeof(int32_t) * (T + 1)));
    int32_t *list = W.big_list.as<int32_t>();
    int32_t *len = list + T;
    HIP_TRY(gs::launch_zero_i32(len, 1, s));
    if (K->big_only) {
        HIP_TRY(W.count.ensure(sizeof(int32_t) * std::max<int64_t>(T, 1)));
        HIP_TRY(gs::launch_fill_i32(W.count.as<int32_t>(), T, 65, s));
    }
    HIP_TRY(gs::launch_collect_over(W.count.as<int32_t>(), T, 64, list, len, s));
    int32_t f = 0;
    HIP_TRY(hipMemcpyAsync(&f, len, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *F = f;
    return GS_OK;
}

// Heap capacity of the large search: the list bound, or for a radius
// search whose bound exceeds the cap, the largest in-radius count among
// the listed targets (capped at kBigCap + 1: a target with more points
// than the cap is refused).
int32_t big_capacity(gs_kriging *K, KWork &W, const double *d_targets, int64_t F, hipStream_t s,
                     int64_t *kcap) {
    if (K->big_only) {
        *kcap = K->stride;
        return GS_OK;
    }
    if (K->big_req <= kBigCap || !std::isfinite(K->r2)) {
        *kcap = std::min<int64_t>(K->big_req, kBigCap + 1);
        return GS_OK;
    }
    gs::KnnParams p;
    std::memset(&p, 0, sizeof(p));
    p.list = W.big_list.as<int32_t>();
    p.tcoords = K->tcoords.as<double>();
    p.tidx = K->tidx.as<uint32_t>();
    p.n = K->n;
    p.targets = d_targets;
    p.T = F;
    p.r2 = K->r2;
    p.xf = K->xf;
    HIP_TRY(W.big_cnt.ensure(sizeof(int32_t) * (F + 1)));
    int32_t *cnt = W.big_cnt.as<int32_t>();
    HIP_TRY(gs::launch_radius_count(K->D, p, cnt, s));
    HIP_TRY(gs::launch_max_i32(cnt, F, cnt + F, s));
    int32_t mx = 0;
    HIP_TRY(hipMemcpyAsync(&mx, cnt + F, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *kcap = std::max<int64_t>(1, std::min<int64_t>(mx, kBigCap + 1));
    return GS_OK;
}

struct BigChunking {
    int64_t stride, slots, chunk, heap_blocks;
};

BigChunking big_chunking(const gs_kriging *K, int64_t F, int64_t kcap) {
    BigChunking c;
    c.stride = K->big_only ? K->stride : kcap;
    c.slots = K->per_oct > 0 ? ((int64_t)K->per_oct << K->D) : kcap;
    const int64_t budget = (int64_t)256 << 20;
    c.chunk = std::min<int64_t>(F, std::max<int64_t>(64, budget / (4 * c.stride)));
    c.heap_blocks = std::min<int64_t>((c.chunk + 63) / 64,
                                      std::max<int64_t>(1, budget / (c.slots * 64 * 12)));
    return c;
}

// The exact kd-tree replay of list[0..Fc) into W.big_nbr / W.big_cnt (by
// list position), heaps in device memory.
int32_t big_search(gs_kriging *K, KWork &W, const double *d_targets, const int32_t *list,
                   int64_t Fc, int64_t kcap, const BigChunking &c, hipStream_t s) {
    HIP_TRY(W.big_nbr.ensure(sizeof(int32_t) * c.chunk * c.stride));
    HIP_TRY(W.big_cnt.ensure(sizeof(int32_t) * (c.chunk + 1)));
    HIP_TRY(W.big_heap_d.ensure(sizeof(double) * c.heap_blocks * c.slots * 64));
    HIP_TRY(W.big_heap_ix.ensure(sizeof(uint32_t) * c.heap_blocks * c.slots * 64));
    gs::KnnParams p;
    std::memset(&p, 0, sizeof(p));
    p.list = list;
    p.tcoords = K->tcoords.as<double>();
    p.tidx = K->tidx.as<uint32_t>();
    p.n = K->n;
    p.targets = d_targets;
    p.T = Fc;
    p.k = (int32_t)kcap;
    p.per_octant = K->per_oct;
    p.max_nb = (int32_t)K->stride;
    p.depth = K->depth;
    p.r2 = K->r2;
    p.xf = K->xf;
    p.out_idx = W.big_nbr.as<int32_t>();
    p.out_count = W.big_cnt.as<int32_t>();
    p.stride = (int32_t)c.stride;
    p.gheap_d = W.big_heap_d.as<double>();
    p.gheap_ix = W.big_heap_ix.as<uint32_t>();
    p.heap_blocks = c.heap_blocks;
    p.out_by_pos = 1;
    HIP_TRY(gs::launch_knn(K->D, p, s));
    return GS_OK;
}

// Targets whose neighbourhood exceeds 64 points: the exact kd-tree replay
// with device-memory heaps, then one workgroup per target solving the
// reference's bordered system by LU (gs_big.hip), in chunks that keep the
// lists, heaps and factorization scratch within bounded device memory.
int32_t run_big(gs_kriging *K, KWork &W, const double *d_targets, int64_t T, double *d_value,
                double *d_var, double *d_lag, int32_t *d_status, hipStream_t s, const Extra &x) {
    int64_t F = 0;
    int32_t st = big_targets(K, W, T, s, &F);
    if (st != GS_OK || F == 0) return st;
    int64_t kcap = 0;
    st = big_capacity(K, W, d_targets, F, s, &kcap);
    if (st != GS_OK) return st;
    const BigChunking c = big_chunking(K, F, kcap);
    const int64_t nmax = std::min<int64_t>(c.stride, kBigCap) + K->M;
    gs::BigParams bp;
    std::memset(&bp, 0, sizeof(bp));
    bp.S = solve_params(K, W, d_targets, T, d_value, d_var, d_lag, d_status, x);
    bp.stride = (int32_t)c.stride;
    bp.M = K->M;
    bp.nv = K->nv;
    bp.cap = (int32_t)kBigCap;
    bp.slot = nmax * nmax + 2 * nmax;
    const int64_t sblocks = std::max<int64_t>(1, ((int64_t)512 << 20) / (8 * bp.slot));
    for (int64_t c0 = 0; c0 < F; c0 += c.chunk) {
        const int64_t Fc = std::min<int64_t>(c.chunk, F - c0);
        const int32_t *list = W.big_list.as<int32_t>() + c0;
        st = big_search(K, W, d_targets, list, Fc, kcap, c, s);
        if (st != GS_OK) return st;
        const int64_t blocks = std::min<int64_t>(Fc, sblocks);
        HIP_TRY(W.big_scratch.ensure(sizeof(double) * blocks * bp.slot));
        bp.list = list;
        bp.F = Fc;
        bp.nbr = W.big_nbr.as<int32_t>();
        bp.count = W.big_cnt.as<int32_t>();
        bp.scratch = W.big_scratch.as<double>();
        HIP_TRY(gs::launch_big_solve(K->D, K->lean, bp, blocks, s));
        g_solve_kernel = "big_solve_kernel";
    }
    return GS_OK;
}

// Scratch bytes of one gs_kriging_predict_many_noalloc call: the workspace
// buffers run_neighbors / run_solves request (no side stream), each rounded
// to the arena's 256-byte units.
size_t noalloc_bytes(const gs_kriging *K, int64_t T, bool r32) {
    auto al = [](size_t b) { return ((b ? b : 8) + 255) / 256 * 256; };
    const size_t t = (size_t)std::max<int64_t>(T, 1);
    size_t n = al(4 * t * (size_t)stride_of(K)) + al(4 * t);      // lists, counts
    if (K->has_tree)
        n += 2 * al(4 * t) + al(gs::tile_sort_workspace(T, ncells_of(K))) + al(4 * (t + 4));
    if (K->per_oct > 0 && K->nb_bound > 16) n += al(4 * 3 * t);  // size buckets
    n += al(4 * 8);                                              // list lengths
    if (r32) n += al(4 * (t + 1));                               // FP64 fallback list
    return n;
}

// The device path of one predict call on workspace W: an external-drift
// predictor fails every target (predict without covariates,
// kriging.rs:494-501, mapped to NaN by predict_many); otherwise run_predict.
int32_t predict_device(gs_kriging *k, KWork &W, const double *d_targets, int64_t T,
                       int32_t precision, double *d_value, double *d_variance, double *d_lagrange,
                       int32_t *d_status, hipStream_t s, bool timed, bool allow_side) {
    if (k->nv > 0) {
        const double nan = std::numeric_limits<double>::quiet_NaN();
        HIP_TRY(gs::launch_fill_f64(d_value, T, nan, s));
        HIP_TRY(gs::launch_fill_f64(d_variance, T, nan, s));
        if (d_lagrange) HIP_TRY(gs::launch_fill_f64(d_lagrange, T, nan, s));
        if (d_status) HIP_TRY(gs::launch_fill_i32(d_status, T, GS_ERR_INVALID_PARAMETER, s));
        return GS_OK;
    }
    return run_predict(k, W, d_targets, T, precision, d_value, d_variance, d_lagrange, d_status, s,
                       Extra(), timed, allow_side);
}

}  // namespace

extern "C" {

int32_t gs_kriging_predict_many_device(gs_kriging *k, const double *d_targets, int64_t T,
                                       int32_t precision, double *d_value, double *d_variance,
                                       double *d_lagrange, int32_t *d_status, void *stream) {
    if (!k) return fail(GS_ERR_INVALID_PARAMETER, "handle is NULL");
    if (T < 0) return fail(GS_ERR_INVALID_PARAMETER, "T < 0");
    if (T > 0 && (!d_targets || !d_value || !d_variance))
        return fail(GS_ERR_INVALID_PARAMETER, "NULL buffer");
    std::lock_guard<std::mutex> lk(k->mu);
    HIP_TRY(hipSetDevice(k->ctx->device));
    return predict_device(k, k->w, d_targets, T, precision, d_value, d_variance, d_lagrange,
                          d_status, (hipStream_t)stream, true, true);
}

// Same as gs_kriging_predict_many_device without any device allocation: the
// workspace comes from the fixed per-device scratch, and the kd-tree replay
// stays on the caller's stream.
int32_t gs_kriging_predict_many_noalloc(gs_kriging *k, const double *d_targets, int64_t T,
                                        int32_t precision, double *d_value, double *d_variance,
                                        double *d_lagrange, int32_t *d_status, void *stream) {
    if (!k) return fail(GS_ERR_INVALID_PARAMETER, "handle is NULL");
    if (T < 0) return fail(GS_ERR_INVALID_PARAMETER, "T < 0");
    if (T > 0 && (!d_targets || !d_value || !d_variance))
        return fail(GS_ERR_INVALID_PARAMETER, "NULL buffer");
    if (precision != GS_PREC_FP64 && precision != GS_PREC_FP32_REFINED)
        return fail(GS_ERR_INVALID_PARAMETER, "unknown precision");
    if (k->global || k->big_req > 0)
        return fail(GS_ERR_INVALID_PARAMETER,
                    "unsupported without allocation: global neighbourhoods and neighbourhoods "
                    "above 64 points need scratch that grows with the data; use "
                    "gs_kriging_predict_many_device");
    if (T == 0) return GS_OK;
    std::lock_guard<std::mutex> lk(k->mu);
    HIP_TRY(hipSetDevice(k->ctx->device));
    void *base = scratch_of(k->ctx->device);
    if (!base) return fail(GS_ERR_DEVICE, "scratch not set up for this device");
    const bool r32 = precision == GS_PREC_FP32_REFINED && k->lean && k->M <= 1;
    const size_t need = noalloc_bytes(k, T, r32);
    if (need > kScratchBytes)
        return fail(GS_ERR_INVALID_PARAMETER,
                    "too many targets for one call without allocation: " + std::to_string(T) +
                        " targets need " + std::to_string(need) + " bytes of scratch, the fixed "
                        "scratch holds " + std::to_string(kScratchBytes) + "; split the targets");
    Arena a;
    a.base = static_cast<char *>(base);
    a.cap = kScratchBytes;
    KWork W;
    W.use_arena(&a);
    return predict_device(k, W, d_targets, T, precision, d_value, d_variance, d_lagrange, d_status,
                          (hipStream_t)stream, false, false);
}

}  // extern "C"

// A predict call captured as a HIP graph over fixed buffers, with a
// workspace of its own (so graph launches and the predictor's other calls
// do not share scratch).  The predictor's resident data is read, never
// written, by the graph.
struct gs_graph {
    int device = 0;
    KWork w;
    hipGraphExec_t exec = nullptr;
    ~gs_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

extern "C" {

// Capture: one eager pass on a private stream sizes every workspace buffer
// and creates the workspace's side stream and events, then the same
// launches are recorded by stream capture, which therefore allocates
// nothing.  Every launch of the path is a kernel (counters are cleared by
// fill kernels, the target order and the selection of unflagged targets are
// scans made of kernels), and the path reads nothing back to the host, so
// the graph holds kernel nodes only.  For large calls the kd-tree replay of
// flagged targets forks onto the side stream and joins back through the
// workspace's events, which capture records as a parallel branch.
int32_t gs_kriging_graph_create(gs_kriging *k, const double *d_targets, int64_t T,
                                int32_t precision, double *d_value, double *d_variance,
                                double *d_lagrange, int32_t *d_status, gs_graph **out) {
    if (!out) return fail(GS_ERR_INVALID_PARAMETER, "out is NULL");
    *out = nullptr;
    if (!k) return fail(GS_ERR_INVALID_PARAMETER, "handle is NULL");
    if (T <= 0) return fail(GS_ERR_INVALID_PARAMETER, "a graph needs T >= 1 targets");
    if (!d_targets || !d_value || !d_variance) return fail(GS_ERR_INVALID_PARAMETER, "NULL buffer");
    if (precision != GS_PREC_FP64 && precision != GS_PREC_FP32_REFINED)
        return fail(GS_ERR_INVALID_PARAMETER, "unknown precision");
    if (k->big_req > 0)
        return fail(GS_ERR_INVALID_PARAMETER,
                    "unsupported in a graph: neighbourhoods above 64 points size their work from "
                    "list lengths read back to the host; use gs_kriging_predict_many_device");
    std::lock_guard<std::mutex> lk(k->mu);
    HIP_TRY(hipSetDevice(k->ctx->device));
    std::unique_ptr<gs_graph> g(new (std::nothrow) gs_graph());
    if (!g) return fail(GS_ERR_OUT_OF_MEMORY, "graph handle");
    g->device = k->ctx->device;
    struct OwnStream {
        hipStream_t s = nullptr;
        ~OwnStream() {
            if (s) (void)hipStreamDestroy(s);
        }
    } os;
    HIP_TRY(hipStreamCreateWithFlags(&os.s, hipStreamNonBlocking));
    int32_t st = predict_device(k, g->w, d_targets, T, precision, d_value, d_variance, d_lagrange,
                                d_status, os.s, false, true);
    if (st != GS_OK) return st;
    HIP_TRY(hipStreamSynchronize(os.s));
    HIP_TRY(hipStreamBeginCapture(os.s, hipStreamCaptureModeThreadLocal));
    st = predict_device(k, g->w, d_targets, T, precision, d_value, d_variance, d_lagrange, d_status,
                        os.s, false, true);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(os.s, &graph);   // ends capture on every path
    if (st != GS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    HIP_TRY(ec);
    const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(ei);
    *out = g.release();
    return GS_OK;
}

int32_t gs_graph_launch(gs_graph *g, void *stream) {
    if (!g) return fail(GS_ERR_INVALID_PARAMETER, "graph is NULL");
    HIP_TRY(hipSetDevice(g->device));
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return GS_OK;
}

void gs_graph_free(gs_graph *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    delete g;
}

}  // extern "C"

namespace {

// Every target failed with `code` (a per-target error the reference's
// predict_many / predict_block_grid map to NaN).
int32_t fail_all_host(int64_t T, double *value, double *variance, double *lagrange,
                      int32_t *status, int32_t code) {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (int64_t t = 0; t < T; t++) {
        value[t] = nan;
        variance[t] = nan;
        if (lagrange) lagrange[t] = nan;
        if (status) status[t] = code;
    }
    return GS_OK;
}

// Host-buffer predict: uploads the targets (and the per-target covariates
// when tdrift is given), runs the pass, downloads the estimates.
int32_t predict_host(gs_kriging *k, KWork &W, const double *targets, int64_t T, int32_t precision,
                     double *value, double *variance, double *lagrange, int32_t *status,
                     const double *tdrift, Extra x, hipStream_t s = nullptr, bool timed = true) {
    if (!s) s = k->ctx->stream;
    HIP_TRY(W.h_targets.ensure(sizeof(double) * T * k->D));
    HIP_TRY(W.h_value.ensure(sizeof(double) * T));
    HIP_TRY(W.h_var.ensure(sizeof(double) * T));
    HIP_TRY(W.h_lag.ensure(sizeof(double) * T));
    HIP_TRY(W.h_status.ensure(sizeof(int32_t) * T));
    HIP_TRY(hipMemcpyAsync(W.h_targets.p, targets, sizeof(double) * T * k->D, hipMemcpyHostToDevice, s));
    if (tdrift) {
        HIP_TRY(W.h_tdrift.ensure(sizeof(double) * T * k->nv));
        HIP_TRY(hipMemcpyAsync(W.h_tdrift.p, tdrift, sizeof(double) * T * k->nv,
                               hipMemcpyHostToDevice, s));
        x.tdrift = W.h_tdrift.as<double>();
    }
    int32_t st = run_predict(k, W, W.h_targets.as<double>(), T, precision, W.h_value.as<double>(),
                             W.h_var.as<double>(), W.h_lag.as<double>(), W.h_status.as<int32_t>(),
                             s, x, timed);
    if (st != GS_OK) return st;
    HIP_TRY(hipMemcpyAsync(value, W.h_value.p, sizeof(double) * T, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(variance, W.h_var.p, sizeof(double) * T, hipMemcpyDeviceToHost, s));
    if (lagrange) HIP_TRY(hipMemcpyAsync(lagrange, W.h_lag.p, sizeof(double) * T, hipMemcpyDeviceToHost, s));
    if (status) HIP_TRY(hipMemcpyAsync(status, W.h_status.p, sizeof(int32_t) * T, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GS_OK;
}

}  // namespace

extern "C" {

int32_t gs_kriging_predict_many(gs_kriging *k, const double *targets, int64_t T, int32_t precision,
                                double *value, double *variance, double *lagrange, int32_t *status) {
    if (!k) return fail(GS_ERR_INVALID_PARAMETER, "handle is NULL");
    if (T < 0) return fail(GS_ERR_INVALID_PARAMETER, "T < 0");
    if (T == 0) return GS_OK;
    if (!targets || !value || !variance) return fail(GS_ERR_INVALID_PARAMETER, "NULL buffer");
    if (precision != GS_PREC_FP64 && precision != GS_PREC_FP32_REFINED)
        return fail(GS_ERR_INVALID_PARAMETER, "unknown precision");
    std::lock_guard<std::mutex> lk(k->mu);
    HIP_TRY(hipSetDevice(k->ctx->device));
    if (k->nv > 0)   // predict needs predict_with_drift here (kriging.rs:494-501)
        return fail_all_host(T, value, variance, lagrange, status, GS_ERR_INVALID_PARAMETER);
    return predict_host(k, k->w, targets, T, precision, value, variance, lagrange, status, nullptr,
                        Extra());
}

int32_t gs_kriging_predict_many_with_drift(gs_kriging *k, const double *targets, int64_t T,
                                           const double *drifts, int32_t precision, double *value,
                                           double *variance, double *lagrange, int32_t *status) {
    if (!k) return fail(GS_ERR_INVALID_PARAMETER, "handle is NULL");
    if (T < 0) return fail(GS_ERR_INVALID_PARAMETER, "T < 0");
    if (T == 0) return GS_OK;
    if (!targets || !value || !variance) return fail(GS_ERR_INVALID_PARAMETER, "NULL buffer");
    if (precision != GS_PREC_FP64 && precision != GS_PREC_FP32_REFINED)
        return fail(GS_ERR_INVALID_PARAMETER, "unknown precision");
    std::lock_guard<std::mutex> lk(k->mu);
    HIP_TRY(hipSetDevice(k->ctx->device));
    // predict_with_drift on another method fails per target (kriging.rs:507-512)
    if (k->nv == 0) return fail_all_host(T, value, variance, lagrange, status, GS_ERR_INVALID_PARAMETER);
    if (!drifts) return fail(GS_ERR_DIMENSION_MISMATCH, "no drift rows for the targets");
    return predict_host(k, k->w, targets, T, precision, value, variance, lagrange, status, drifts,
                        Extra());
}

int32_t gs_kriging_predict_block_many(gs_kriging *k, const double *centers, int64_t T,
                                      const double *offsets, int32_t n_off, int32_t precision,
                                      double *value, double *variance, double *lagrange,
                                      int32_t *status) {
    if (!k) return fail(GS_ERR_INVALID_PARAMETER, "handle is NULL");
    if (T < 0) return fail(GS_ERR_INVALID_PARAMETER, "T < 0");
    if (T == 0) return GS_OK;
    if (!centers || !value || !variance) return fail(GS_ERR_INVALID_PARAMETER, "NULL buffer");
    if (precision != GS_PREC_FP64 && precision != GS_PREC_FP32_REFINED)
        return fail(GS_ERR_INVALID_PARAMETER, "unknown precision");
    if (n_off < 0) return fail(GS_ERR_INVALID_PARAMETER, "n_off < 0");
    std::lock_guard<std::mutex> lk(k->mu);
    HIP_TRY(hipSetDevice(k->ctx->device));
    // predict_block's own errors (kriging.rs:679-697) fail every target
    if (k->cfg.method != GS_SIMPLE && k->cfg.method != GS_ORDINARY)
        return fail_all_host(T, value, variance, lagrange, status, GS_ERR_INVALID_PARAMETER);
    if (n_off == 0 || !offsets)
        return fail_all_host(T, value, variance, lagrange, status, GS_ERR_INVALID_PARAMETER);
    for (int64_t i = 0; i < (int64_t)n_off * k->D; i++)
        if (!std::isfinite(offsets[i])) return fail(GS_ERR_INVALID_PARAMETER, "non-finite block offset");
    hipStream_t s = k->ctx->stream;
    const size_t ob = sizeof(double) * (size_t)n_off * k->D;
    HIP_TRY(k->w.blk.ensure(ob + 64));
    double *d_offs = k->w.blk.as<double>();
    double *d_cbb = reinterpret_cast<double *>(static_cast<char *>(k->w.blk.p) + ob);
    HIP_TRY(hipMemcpyAsync(d_offs, offsets, ob, hipMemcpyHostToDevice, s));
    HIP_TRY(gs::launch_block_cbb(k->D, k->lean, k->dmodel.as<DevModel>(), d_offs, n_off, d_cbb, s));
    Extra x;
    x.offsets = d_offs;
    x.n_off = n_off;
    x.cbb = d_cbb;
    return predict_host(k, k->w, centers, T, precision, value, variance, lagrange, status, nullptr,
                        x);
}

int32_t gs_kriging_neighbors(gs_kriging *k, const double *targets, int64_t T, int64_t max_out,
                             int32_t *idx, int32_t *count) {
    if (!k) return fail(GS_ERR_INVALID_PARAMETER, "handle is NULL");
    if (T < 0) return fail(GS_ERR_INVALID_PARAMETER, "T < 0");
    if (T == 0) return GS_OK;
    if (!targets || !idx || !count) return fail(GS_ERR_INVALID_PARAMETER, "NULL buffer");
    if (max_out < gs_kriging_max_neighbors(k))
        return fail(GS_ERR_INVALID_PARAMETER, "max_out smaller than the neighbourhood bound");
    if (k->global) {   // every datum, in index order (kriging.rs:441-444)
        for (int64_t t = 0; t < T; t++) {
            for (int64_t j = 0; j < max_out; j++) idx[t * max_out + j] = j < k->n ? (int32_t)j : -1;
            count[t] = (int32_t)k->n;
        }
        return GS_OK;
    }
    std::lock_guard<std::mutex> lk(k->mu);
    HIP_TRY(hipSetDevice(k->ctx->device));
    hipStream_t s = k->ctx->stream;
    HIP_TRY(k->w.h_targets.ensure(sizeof(double) * T * k->D));
    HIP_TRY(hipMemcpyAsync(k->w.h_targets.p, targets, sizeof(double) * T * k->D, hipMemcpyHostToDevice, s));
    const double *d_targets = k->w.h_targets.as<double>();
    if (!k->big_only) {
        int32_t st = run_neighbors(k, k->w, d_targets, T, s);
        if (st != GS_OK) return st;
        const int32_t stride = stride_of(k);
        std::vector<int32_t> tmp((size_t)(T * stride));
        HIP_TRY(hipMemcpyAsync(tmp.data(), k->w.nbr.p, sizeof(int32_t) * T * stride, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(count, k->w.count.p, sizeof(int32_t) * T, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int64_t t = 0; t < T; t++)
            for (int64_t j = 0; j < max_out; j++)
                idx[t * max_out + j] = j < stride ? tmp[(size_t)(t * stride + j)] : -1;
    }
    if (k->big_req == 0) return GS_OK;
    // lists above 64 points: the large search for the targets that need it
    int64_t F = 0;
    int32_t st = big_targets(k, k->w, T, s, &F);
    if (st != GS_OK || F == 0) return st;
    int64_t kcap = 0;
    st = big_capacity(k, k->w, d_targets, F, s, &kcap);
    if (st != GS_OK) return st;
    const BigChunking c = big_chunking(k, F, kcap);
    std::vector<int32_t> list((size_t)F), lists, cnts;
    HIP_TRY(hipMemcpyAsync(list.data(), k->w.big_list.p, sizeof(int32_t) * F, hipMemcpyDeviceToHost, s));
    for (int64_t c0 = 0; c0 < F; c0 += c.chunk) {
        const int64_t Fc = std::min<int64_t>(c.chunk, F - c0);
        st = big_search(k, k->w, d_targets, k->w.big_list.as<int32_t>() + c0, Fc, kcap, c, s);
        if (st != GS_OK) return st;
        lists.resize((size_t)(Fc * c.stride));
        cnts.resize((size_t)Fc);
        HIP_TRY(hipMemcpyAsync(lists.data(), k->w.big_nbr.p, sizeof(int32_t) * Fc * c.stride,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(cnts.data(), k->w.big_cnt.p, sizeof(int32_t) * Fc, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int64_t q = 0; q < Fc; q++) {
            const int64_t t = list[(size_t)(c0 + q)];
            count[t] = cnts[(size_t)q];
            for (int64_t j = 0; j < max_out; j++)
                idx[t * max_out + j] = j < c.stride ? lists[(size_t)(q * c.stride + j)] : -1;
        }
    }
    return GS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Cross-validation (validation.rs:136-400): each fold builds a predictor on
// the points outside it and predicts its members, as the reference does
// (PointSet::excluding / the training subset, then Kriging::new and predict).
// ---------------------------------------------------------------------------
namespace {

// rng.rs Rng: splitmix64 seeding, xoshiro256++ and Fisher-Yates shuffle
// (rng.rs:8-52, 76-82); the fold assignment depends on the exact stream.
struct FoldRng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &st) {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    explicit FoldRng(uint64_t seed) {
        uint64_t sm = seed;
        for (int i = 0; i < 4; i++) s[i] = splitmix(sm);
    }
    uint64_t next() {
        const uint64_t r = rotl(s[0] + s[3], 23) + s[0];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    template <class T>
    void shuffle(std::vector<T> &xs) {
        for (size_t i = xs.size(); i-- > 1;) {
            const size_t j = (size_t)(next() % (uint64_t)(i + 1));
            std::swap(xs[i], xs[j]);
        }
    }
};

// cv_with_folds (validation.rs:214-256): folds in index order, empty folds
// skipped; the first failing fold (and member) decides the error.
int32_t cv_with_folds(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                      int64_t n, const gs_model *model, const gs_kconfig *cfg, int32_t nv,
                      const double *drift, const std::vector<int64_t> &fold_of,
                      const std::vector<std::vector<int64_t>> &members, double *predicted,
                      double *variance) {
    std::vector<double> tc, tv, td, mc, md, ov, ovar, olag;
    std::vector<int32_t> ost;
    for (size_t f = 0; f < members.size(); f++) {
        const auto &mem = members[f];
        if (mem.empty()) continue;
        tc.clear();
        tv.clear();
        td.clear();
        for (int64_t i = 0; i < n; i++) {
            if (fold_of[(size_t)i] == (int64_t)f) continue;
            for (int d = 0; d < dim; d++) tc.push_back(coords[i * dim + d]);
            tv.push_back(values[i]);
            for (int k = 0; k < nv; k++) td.push_back(drift[i * nv + k]);
        }
        const int64_t nt = (int64_t)tv.size();
        if (nt == 0) return fail(GS_ERR_INSUFFICIENT_DATA, "cross-validation fold has no training points");
        gs_kriging *K = nullptr;
        int64_t dup[2];
        int32_t st = kriging_build(ctx, dim, tc.data(), tv.data(), nt, model, cfg, nullptr, nv,
                                   nv > 0 ? td.data() : nullptr, &K, dup);
        if (st != GS_OK) return st;
        std::unique_ptr<gs_kriging> own(K);
        const int64_t T = (int64_t)mem.size();
        mc.resize((size_t)(T * dim));
        md.resize((size_t)(T * std::max(nv, 1)));
        for (int64_t j = 0; j < T; j++) {
            for (int d = 0; d < dim; d++) mc[(size_t)(j * dim + d)] = coords[mem[(size_t)j] * dim + d];
            for (int k = 0; k < nv; k++) md[(size_t)(j * nv + k)] = drift[mem[(size_t)j] * nv + k];
        }
        ov.resize((size_t)T);
        ovar.resize((size_t)T);
        olag.resize((size_t)T);
        ost.resize((size_t)T);
        {
            std::lock_guard<std::mutex> lk(K->mu);
            HIP_TRY(hipSetDevice(ctx->device));
            st = predict_host(K, K->w, mc.data(), T, GS_PREC_FP64, ov.data(), ovar.data(), olag.data(),
                              ost.data(), nv > 0 ? md.data() : nullptr, Extra());
        }
        if (st != GS_OK) return st;
        for (int64_t j = 0; j < T; j++) {
            if (ost[(size_t)j] != GS_OK)
                return fail(ost[(size_t)j], "cross-validation: prediction at point " +
                                                std::to_string(mem[(size_t)j]) + " failed");
            predicted[mem[(size_t)j]] = ov[(size_t)j];
            variance[mem[(size_t)j]] = ovar[(size_t)j];
        }
    }
    return GS_OK;
}

int32_t cv_args(gs_ctx *ctx, int32_t dim, const double *coords, const double *values, int64_t n,
                const gs_model *model, const gs_kconfig *cfg, double *predicted,
                double *variance) {
    gs_kriging *dummy = nullptr;
    int32_t st = check_pointset(ctx, dim, coords, values, n, model, cfg, &dummy);
    if (st != GS_OK) return st;
    if (!predicted || !variance) return fail(GS_ERR_INVALID_PARAMETER, "NULL output buffer");
    if (n < 3) return fail(GS_ERR_INSUFFICIENT_DATA, "cross-validation requires at least 3 points");
    return GS_OK;
}

}  // namespace

extern "C" {

int32_t gs_cv_leave_one_out(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                            int64_t n, const gs_model *model, const gs_kconfig *config,
                            double *predicted, double *variance) {
    int32_t st = cv_args(ctx, dim, coords, values, n, model, config, predicted, variance);
    if (st != GS_OK) return st;
    if (config->method == GS_EXTERNAL_DRIFT)
        return fail(GS_ERR_INVALID_PARAMETER, "external drift needs gs_cv_leave_one_out_with_drift");
    std::vector<int64_t> fold_of((size_t)n);
    std::vector<std::vector<int64_t>> members((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        fold_of[(size_t)i] = i;
        members[(size_t)i].push_back(i);
    }
    return cv_with_folds(ctx, dim, coords, values, n, model, config, 0, nullptr, fold_of, members,
                         predicted, variance);
}

int32_t gs_cv_leave_one_out_with_drift(gs_ctx *ctx, int32_t dim, const double *coords,
                                       const double *values, int64_t n, const gs_model *model,
                                       const gs_kconfig *config, int32_t n_vars,
                                       const double *drift_data, double *predicted,
                                       double *variance) {
    int32_t st = cv_args(ctx, dim, coords, values, n, model, config, predicted, variance);
    if (st != GS_OK) return st;
    if (!drift_data) return fail(GS_ERR_DIMENSION_MISMATCH, "no drift rows for the data points");
    // with_external_drift's checks run per fold in the reference; they do
    // not depend on the fold
    if (config->method != GS_EXTERNAL_DRIFT)
        return fail(GS_ERR_INVALID_PARAMETER, "with_external_drift requires KrigingMethod::ExternalDrift");
    if (n_vars <= 0) return fail(GS_ERR_INVALID_PARAMETER, "external drift requires at least one covariate");
    for (int64_t i = 0; i < n * n_vars; i++)
        if (!std::isfinite(drift_data[i]))
            return fail(GS_ERR_INVALID_PARAMETER, "non-finite drift value");
    if (n_vars > 3)
        return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: more than 3 external covariates");
    std::vector<int64_t> fold_of((size_t)n);
    std::vector<std::vector<int64_t>> members((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        fold_of[(size_t)i] = i;
        members[(size_t)i].push_back(i);
    }
    return cv_with_folds(ctx, dim, coords, values, n, model, config, n_vars, drift_data, fold_of,
                         members, predicted, variance);
}

int32_t gs_cv_k_fold(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                     int64_t n, const gs_model *model, const gs_kconfig *config, int64_t k,
                     uint64_t seed, double *predicted, double *variance) {
    int32_t st = cv_args(ctx, dim, coords, values, n, model, config, predicted, variance);
    if (st != GS_OK) return st;
    if (k < 2 || k > n) return fail(GS_ERR_INVALID_PARAMETER, "k-fold requires 2 <= k <= n");
    if (config->method == GS_EXTERNAL_DRIFT)
        return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: k-fold with external drift");
    // fold_assignment (validation.rs:199-212)
    std::vector<int64_t> order((size_t)n);
    for (int64_t i = 0; i < n; i++) order[(size_t)i] = i;
    FoldRng(seed).shuffle(order);
    std::vector<int64_t> fold_of((size_t)n);
    std::vector<std::vector<int64_t>> members((size_t)k);
    for (int64_t j = 0; j < n; j++) {
        const int64_t i = order[(size_t)j];
        fold_of[(size_t)i] = j % k;
        members[(size_t)(j % k)].push_back(i);
    }
    return cv_with_folds(ctx, dim, coords, values, n, model, config, 0, nullptr, fold_of, members,
                         predicted, variance);
}

int32_t gs_cv_block(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                    int64_t n, const gs_model *model, const gs_kconfig *config,
                    const int64_t *blocks_per_dim, double *predicted, double *variance) {
    int32_t st = cv_args(ctx, dim, coords, values, n, model, config, predicted, variance);
    if (st != GS_OK) return st;
    if (!blocks_per_dim) return fail(GS_ERR_INVALID_PARAMETER, "blocks_per_dim is NULL");
    int64_t nb = 1;
    for (int d = 0; d < dim; d++) {
        if (blocks_per_dim[d] <= 0)
            return fail(GS_ERR_INVALID_PARAMETER, "blocks_per_dim entries must be at least 1");
        nb *= blocks_per_dim[d];
    }
    if (nb < 2) return fail(GS_ERR_INVALID_PARAMETER, "block CV requires at least 2 blocks total");
    if (config->method == GS_EXTERNAL_DRIFT)
        return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: block CV with external drift");
    // spatial_block_assignment (validation.rs:294-323)
    double mn[3], mx[3];
    for (int d = 0; d < dim; d++) mn[d] = mx[d] = coords[d];
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < dim; d++) {
            mn[d] = std::fmin(mn[d], coords[i * dim + d]);
            mx[d] = std::fmax(mx[d], coords[i * dim + d]);
        }
    std::vector<int64_t> fold_of((size_t)n);
    std::vector<std::vector<int64_t>> members((size_t)nb);
    for (int64_t i = 0; i < n; i++) {
        int64_t idx = 0, stride = 1;
        for (int d = 0; d < dim; d++) {
            const double span = std::fmax(mx[d] - mn[d], std::numeric_limits<double>::min());
            const double frac = std::clamp((coords[i * dim + d] - mn[d]) / span, 0.0, 1.0 - 1e-12);
            // float -> usize cast saturates at 0 and truncates toward zero
            const double scaled = frac * (double)blocks_per_dim[d];
            int64_t bd = scaled > 0.0 ? (int64_t)scaled : 0;
            bd = std::min<int64_t>(bd, blocks_per_dim[d] - 1);
            idx += bd * stride;
            stride *= blocks_per_dim[d];
        }
        fold_of[(size_t)i] = idx;
        members[(size_t)idx].push_back(i);
    }
    return cv_with_folds(ctx, dim, coords, values, n, model, config, 0, nullptr, fold_of, members,
                         predicted, variance);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Vecchia plan: maxmin ordering + nearest predecessors (vecchia.rs:144-282)
// ---------------------------------------------------------------------------
namespace {

double dist2(const double *a, const double *b, int D) {   // vecchia.rs:95-102
    double s = 0.0;
    for (int d = 0; d < D; d++) {
        double v = a[d] - b[d];
        s += v * v;
    }
    return s;
}

// Bucket grid with the reference's sizing rule and shell-walk order
// (search.rs:119-313).  Entries carry their insertion position; a query can
// be restricted to positions < limit, which replays the reference's
// incremental grid exactly (same buckets, same in-bucket order).
struct BGrid {
    int D;
    double origin[3];
    double cell;
    int64_t nn[3], strides[3];
    std::vector<int64_t> start;      // CSR buckets (built once) ...
    std::vector<uint32_t> items;
    std::vector<std::vector<uint32_t>> dyn;   // ... or growable buckets
    const double *pts;               // [pos][D]
    bool csr = false;

    void init(int D_, const double *mn, const double *mx, int64_t expected) {   // search.rs:138-197
        D = D_;
        double ext[3] = {0, 0, 0}, max_extent = 0.0;
        for (int d = 0; d < D; d++) { ext[d] = mx[d] - mn[d]; max_extent = std::fmax(max_extent, ext[d]); }
        double tol = max_extent * 1e-9;
        bool active[3];
        int n_active = 0;
        for (int d = 0; d < D; d++) { active[d] = ext[d] > tol; n_active += active[d]; }
        if (n_active < 1) n_active = 1;
        double volume = 1.0;
        for (int d = 0; d < D; d++) if (active[d]) volume *= ext[d];
        int64_t em = std::max<int64_t>(expected, 1);
        double c = std::fmax(std::pow(volume, 1.0 / (double)n_active), 1e-12) *
                   (1.0 / std::pow((double)em, 1.0 / (double)n_active));
        int64_t cap = std::max<int64_t>(em * 4, 64);
        for (;;) {
            int64_t total = 1;
            bool over = false;
            for (int d = 0; d < D; d++) {
                if (active[d]) {
                    double q = std::ceil(ext[d] / c);
                    int64_t v = q >= 9.2e18 ? INT64_MAX : (q > 0 ? (int64_t)q : 0);
                    nn[d] = std::max<int64_t>(v, 1);
                } else nn[d] = 1;
                if (total > INT64_MAX / nn[d]) { over = true; break; }
                total *= nn[d];
            }
            if (!over && total <= cap) break;
            c *= 2.0;
        }
        int64_t total = 1;
        for (int d = 0; d < D; d++) { strides[d] = total; total *= nn[d]; origin[d] = mn[d]; }
        cell = c;
        dyn.clear();
    }
    int64_t nbuckets() const {
        int64_t t = 1;
        for (int d = 0; d < D; d++) t *= nn[d];
        return t;
    }
    static int64_t sat(double v) {
        if (v != v) return 0;
        if (v >= 9.2233720368547758e18) return INT64_MAX;
        if (v <= -9.2233720368547758e18) return INT64_MIN;
        return (int64_t)v;
    }
    void cell_of(const double *p, int64_t *c) const {   // search.rs:199-206
        for (int d = 0; d < D; d++) {
            int64_t v = sat((p[d] - origin[d]) / cell);
            c[d] = std::min(std::max<int64_t>(v, 0), nn[d] - 1);
        }
    }
    int64_t bucket_of(const double *p) const {
        int64_t c[3];
        cell_of(p, c);
        int64_t b = 0;
        for (int d = 0; d < D; d++) b += c[d] * strides[d];
        return b;
    }
    // All points inserted in position order, stored as CSR.
    void build_all(const double *p, int64_t npos) {
        pts = p;
        int64_t nb = nbuckets();
        start.assign((size_t)nb + 1, 0);
        std::vector<int64_t> bk((size_t)npos);
        for (int64_t i = 0; i < npos; i++) { bk[(size_t)i] = bucket_of(p + i * D); start[(size_t)bk[(size_t)i] + 1]++; }
        for (int64_t b = 0; b < nb; b++) start[(size_t)b + 1] += start[(size_t)b];
        items.resize((size_t)npos);
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < npos; i++) items[(size_t)fill[(size_t)bk[(size_t)i]]++] = (uint32_t)i;
        csr = true;
    }
};

// Bounded max-heap over (d2, pos) with the reference's push rule
// (search.rs:15-30).
struct KHeap {
    std::vector<std::pair<double, uint32_t>> v;
    int64_t k;
    static bool gt(const std::pair<double, uint32_t> &a, const std::pair<double, uint32_t> &b) {
        return a.first > b.first || (a.first == b.first && a.second > b.second);
    }
    void reset(int64_t k_) { k = k_; v.clear(); }
    double worst() const { return (int64_t)v.size() == k ? v[0].first : INFINITY; }
    void push(double d2, uint32_t i) {
        if ((int64_t)v.size() < k) {
            v.push_back({d2, i});
            size_t e = v.size() - 1;
            while (e > 0) {
                size_t p = (e - 1) / 2;
                if (gt(v[e], v[p])) { std::swap(v[e], v[p]); e = p; } else break;
            }
        } else if (d2 < v[0].first) {
            v[0] = {d2, i};
            size_t e = 0, n = v.size();
            for (;;) {
                size_t l = 2 * e + 1, r = l + 1, g = e;
                if (l < n && gt(v[l], v[g])) g = l;
                if (r < n && gt(v[r], v[g])) g = r;
                if (g == e) break;
                std::swap(v[e], v[g]);
                e = g;
            }
        }
    }
};

// BucketGrid::k_nearest (search.rs:221-313) over positions < limit.
void grid_knn(const BGrid &G, const double *target, int64_t k, int64_t limit, KHeap &h,
              std::vector<uint32_t> &out) {
    out.clear();
    h.reset(k);
    if (k == 0 || limit == 0) return;
    const int D = G.D;
    int64_t tc[3];
    G.cell_of(target, tc);
    int64_t max_ring = G.nn[0];
    for (int d = 1; d < D; d++) max_ring = std::max(max_ring, G.nn[d]);
    auto visit = [&](const int64_t *cur) {
        for (int d = 0; d < D; d++) if (cur[d] < 0 || cur[d] >= G.nn[d]) return;
        int64_t b = 0;
        for (int d = 0; d < D; d++) b += cur[d] * G.strides[d];
        for (int64_t q = G.start[(size_t)b]; q < G.start[(size_t)b + 1]; q++) {
            uint32_t i = G.items[(size_t)q];
            if ((int64_t)i >= limit) break;   // positions ascend inside a bucket
            double d2 = 0.0;
            for (int d = 0; d < D; d++) {
                double dd = G.pts[(int64_t)i * D + d] - target[d];
                d2 += dd * dd;
            }
            h.push(d2, i);
        }
    };
    for (int64_t ring = 0; ring <= max_ring; ring++) {
        double rm = std::fmax((double)ring - 1.0, 0.0) * G.cell;
        if (rm * rm > h.worst()) break;
        if (ring == 0) { visit(tc); continue; }
        int64_t r = ring;
        for (int fd = 0; fd < D; fd++) {
            for (int sd = 0; sd < 2; sd++) {
                int64_t side = sd == 0 ? -r : r;
                int64_t lo[3], hi[3];
                bool empty = false;
                for (int d = 0; d < D; d++) {
                    if (d < fd) { lo[d] = tc[d] - r + 1; hi[d] = tc[d] + r - 1; }
                    else if (d == fd) { lo[d] = tc[d] + side; hi[d] = tc[d] + side; }
                    else { lo[d] = tc[d] - r; hi[d] = tc[d] + r; }
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
    std::sort(h.v.begin(), h.v.end());
    for (auto &e : h.v) out.push_back(e.second);
}

// maxmin_order (vecchia.rs:144-215): exact farthest-point traversal, ties
// to the larger index.  The reference refreshes stale heap keys lazily;
// here keys are kept exact instead: when p is selected with key l2 (the
// largest current min-distance), only points closer to p than sqrt(l2) can
// lower their key, so one range query on a grid of all points updates every
// key to the min squared distance to the selected set.  The next point is
// the argmax of (key, index), found through a tournament tree over blocks of
// 64 bucket-ordered slots: a lowered key only costs a block rescan when it
// was its block's maximum.  The order is the same as the lazy heap's; the
// work is ~n log n without the lazy refreshes' 1-NN queries.
//
// The set-up (first point, bucket grid, bucket-ordered coordinates) is
// shared with the GPU traversal, which hands its keys to the host traversal
// when its batches stay small.
struct MaxminSetup {
    int D = 2;
    int64_t n = 0;
    int64_t first = 0;
    uint32_t qf = 0;                 // slot of the first point
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    BGrid G;                         // ~1 point per bucket over all points
    std::vector<double> cc;          // [slot][D]
};

void maxmin_setup(int D, const double *coords, int64_t n, MaxminSetup &S) {
    S.D = D;
    S.n = n;
    double c[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < D; d++) c[d] += coords[i * D + d] / (double)n;
    int64_t first = 0;
    double bd = dist2(coords, c, D);
    for (int64_t i = 1; i < n; i++) {
        double dd = dist2(coords + i * D, c, D);
        if (total_key(dd) < total_key(bd)) { bd = dd; first = i; }
    }
    S.first = first;
    for (int d = 0; d < D; d++) { S.mn[d] = coords[d]; S.mx[d] = coords[d]; }
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < D; d++) {
            S.mn[d] = std::fmin(S.mn[d], coords[i * D + d]);
            S.mx[d] = std::fmax(S.mx[d], coords[i * D + d]);
        }
    S.G.init(D, S.mn, S.mx, n);
    S.G.build_all(coords, n);
    // Everything below lives in bucket order (slot q holds point items[q]),
    // so a range query reads contiguous coordinates and keys.
    const std::vector<uint32_t> &id = S.G.items;
    S.cc.resize((size_t)(n * D));
    for (int64_t q = 0; q < n; q++) {
        if ((int64_t)id[(size_t)q] == first) S.qf = (uint32_t)q;
        for (int d = 0; d < D; d++) S.cc[(size_t)(q * D + d)] = coords[(int64_t)id[(size_t)q] * D + d];
    }
}

// Continues the traversal: `order` holds the points ordered so far and
// key[q] each other slot's squared distance to them (-1 for ordered slots).
// Keys are sums of squares (never -0.0 or NaN), for which total_cmp is the
// plain numeric order.
void maxmin_traverse(MaxminSetup &S, std::vector<double> &key, std::vector<int64_t> &order) {
    const int D = S.D;
    const int64_t n = S.n;
    const BGrid &G = S.G;
    const std::vector<uint32_t> &id = G.items;
    const std::vector<double> &cc = S.cc;
    if ((int64_t)order.size() >= n) return;
    // per block: the slot of its (key, index) maximum, as MaxminKey
    // (vecchia.rs:114-125), with that key and index
    constexpr int64_t kB = 64;
    const int64_t nblk = (n + kB - 1) / kB;
    std::vector<uint32_t> bmax((size_t)nblk), bid((size_t)nblk);
    std::vector<double> bkey((size_t)nblk);
    auto rescan = [&](int64_t b) {
        const int64_t q0 = b * kB, len = std::min<int64_t>(kB, n - q0);
        const double *kk = key.data() + q0;
        double m = kk[0];
        for (int64_t j = 1; j < len; j++) m = kk[j] > m ? kk[j] : m;
        uint32_t best = (uint32_t)q0, bi = 0;
        bool found = false;
        for (int64_t j = 0; j < len; j++)
            if (kk[j] == m && (!found || id[(size_t)(q0 + j)] > bi)) {
                best = (uint32_t)(q0 + j);
                bi = id[(size_t)(q0 + j)];
                found = true;
            }
        bmax[(size_t)b] = best;
        bkey[(size_t)b] = m;
        bid[(size_t)b] = bi;
    };
    for (int64_t b = 0; b < nblk; b++) rescan(b);
    // tournament tree: node i holds the winning block of its range
    int64_t L = 1;
    while (L < nblk) L <<= 1;
    std::vector<int32_t> win((size_t)(2 * L), -1);
    auto better = [&](int32_t a, int32_t b) -> int32_t {
        if (a < 0) return b;
        if (b < 0) return a;
        const double ka = bkey[(size_t)a], kb = bkey[(size_t)b];
        return (kb > ka || (kb == ka && bid[(size_t)b] > bid[(size_t)a])) ? b : a;
    };
    for (int64_t b = 0; b < nblk; b++) win[(size_t)(L + b)] = (int32_t)b;
    for (int64_t i = L - 1; i >= 1; i--) win[(size_t)i] = better(win[(size_t)(2 * i)], win[(size_t)(2 * i + 1)]);
    // block b's maximum went down: replay its path towards the root, up to
    // the first node that neither changes nor was won by b
    auto lowered = [&](int64_t b) {
        rescan(b);
        for (int64_t i = (L + b) >> 1; i >= 1; i >>= 1) {
            const int32_t w = better(win[(size_t)(2 * i)], win[(size_t)(2 * i + 1)]);
            if (w == win[(size_t)i] && w != (int32_t)b) break;
            win[(size_t)i] = w;
        }
    };
    order.reserve((size_t)n);
    for (int64_t sel = (int64_t)order.size(); sel < n; sel++) {
        const int32_t wb = win[1];
        const uint32_t p = bmax[(size_t)wb];
        const double kp = key[p];
        order.push_back((int64_t)id[p]);
        key[p] = -1.0;
        lowered(wb);
        if (sel + 1 == n) break;
        // points whose key can drop lie within sqrt(key[p]) of p (every
        // remaining key is <= key[p]); the box is widened for rounding
        const double *pp = &cc[(size_t)p * D];
        const double rad = std::sqrt(kp) * (1.0 + 1e-9) + 1e-300;
        int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int d = 0; d < D; d++) {
            lo[d] = std::min(std::max<int64_t>(BGrid::sat((pp[d] - rad - G.origin[d]) / G.cell), 0),
                             G.nn[d] - 1);
            hi[d] = std::min(std::max<int64_t>(BGrid::sat((pp[d] + rad - G.origin[d]) / G.cell), 0),
                             G.nn[d] - 1);
        }
        int64_t cur[3] = {lo[0], lo[1], lo[2]};
        for (;;) {
            // one x-run of buckets is one contiguous slot range
            int64_t b0 = 0;
            for (int d = 1; d < D; d++) b0 += cur[d] * G.strides[d];
            const int64_t q0 = G.start[(size_t)(b0 + lo[0])], q1 = G.start[(size_t)(b0 + hi[0]) + 1];
            int64_t blk = -1;
            bool hit = false;   // the block's maximum was lowered
            for (int64_t q = q0; q < q1; q++) {
                if (q / kB != blk) {
                    if (hit) lowered(blk);
                    blk = q / kB;
                    hit = false;
                }
                const double d2 = dist2(&cc[(size_t)(q * D)], pp, D);
                if (d2 < key[(size_t)q]) {
                    key[(size_t)q] = d2;
                    hit = hit || (uint32_t)q == bmax[(size_t)blk];
                }
            }
            if (hit) lowered(blk);
            int d = 1;
            for (; d < D; d++) {
                if (++cur[d] <= hi[d]) break;
                cur[d] = lo[d];
            }
            if (d >= D) break;
        }
    }
}

void maxmin_order(int D, const double *coords, int64_t n, std::vector<int64_t> &order) {
    order.clear();
    if (n == 0) return;
    MaxminSetup S;
    maxmin_setup(D, coords, n, S);
    order.push_back(S.first);
    if (n == 1) return;
    std::vector<double> key((size_t)n);    // min squared distance to the selected set
    for (int64_t q = 0; q < n; q++)
        key[(size_t)q] = dist2(&S.cc[(size_t)(q * D)], coords + S.first * D, D);
    key[S.qf] = -1.0;
    maxmin_traverse(S, key, order);
}

// The same order on the GPU (gs_maxmin.hip): rounds of batched appends,
// driven in chunks; the host finishes the traversal from the device's keys
// when the recent rounds appended few points each (ties of one key value
// make them retry, and lattices make batches small).  info (may be NULL):
// rounds run, rounds that raised the threshold, points ordered on the GPU.
int32_t maxmin_order_gpu(gs_ctx *ctx, int D, const double *coords, int64_t n,
                         std::vector<int64_t> &order, int64_t *info, hipStream_t s) {
    order.clear();
    if (info) { info[0] = 0; info[1] = 0; info[2] = 0; }
    if (n == 0) return GS_OK;
    if (n >= (int64_t)1 << 31) return fail(GS_ERR_INVALID_PARAMETER, "more than 2^31 - 1 points");
    MaxminSetup S;
    maxmin_setup(D, coords, n, S);
    if (n == 1) {
        order.push_back(S.first);
        if (info) info[2] = 1;
        return GS_OK;
    }
    const BGrid &G = S.G;
    HIP_TRY(hipSetDevice(ctx->device));
    const int cap = gs::kMaxminCap;
    DBuf d_cc, d_id, d_start, d_enc, d_cenc, d_cslot, d_bslot, d_brow, d_bbox, d_hist, d_order, d_st;
    HIP_TRY(d_cc.ensure(sizeof(double) * n * D));
    HIP_TRY(d_id.ensure(sizeof(uint32_t) * n));
    HIP_TRY(d_start.ensure(sizeof(int64_t) * G.start.size()));
    HIP_TRY(d_enc.ensure(sizeof(unsigned long long) * n));
    HIP_TRY(d_cenc.ensure(sizeof(unsigned long long) * cap));
    HIP_TRY(d_cslot.ensure(sizeof(uint32_t) * cap));
    HIP_TRY(d_bslot.ensure(sizeof(uint32_t) * cap));
    HIP_TRY(d_brow.ensure(sizeof(int64_t) * (cap + 1)));
    HIP_TRY(d_bbox.ensure(sizeof(int32_t) * 6 * cap));
    HIP_TRY(d_hist.ensure(sizeof(int32_t) * gs::kMaxminBins));
    HIP_TRY(d_order.ensure(sizeof(int64_t) * n));
    HIP_TRY(d_st.ensure(sizeof(gs::MaxminState)));
    HIP_TRY(hipMemcpyAsync(d_cc.p, S.cc.data(), sizeof(double) * n * D, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_id.p, G.items.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_start.p, G.start.data(), sizeof(int64_t) * G.start.size(),
                           hipMemcpyHostToDevice, s));
    HIP_TRY(gs::launch_zero_i32(d_hist.as<int32_t>(), gs::kMaxminBins, s));
    // initial threshold: a power of two above every squared distance
    double diag2 = 0.0;
    for (int d = 0; d < D; d++) diag2 += (S.mx[d] - S.mn[d]) * (S.mx[d] - S.mn[d]);
    uint64_t tb;
    {
        const double t = diag2 * 4.0 + 1e-300;
        std::memcpy(&tb, &t, sizeof(tb));
    }
    gs::MaxminState st0;
    std::memset(&st0, 0, sizeof(st0));
    st0.tau_key = ((tb >> 52) + 1) << 52;
    st0.s = 1;
    HIP_TRY(hipMemcpyAsync(d_st.p, &st0, sizeof(st0), hipMemcpyHostToDevice, s));
    gs::MaxminParams P;
    std::memset(&P, 0, sizeof(P));
    P.n = n;
    P.cc = d_cc.as<double>();
    P.id = d_id.as<uint32_t>();
    P.enc = d_enc.as<unsigned long long>();
    for (int d = 0; d < 3; d++) {
        P.first[d] = d < D ? coords[S.first * D + d] : 0.0;
        P.origin[d] = d < D ? G.origin[d] : 0.0;
        P.nn[d] = d < D ? G.nn[d] : 1;
        P.strides[d] = d < D ? G.strides[d] : 0;
    }
    P.first_slot = S.qf;
    P.cell = G.cell;
    P.start = d_start.as<int64_t>();
    P.cap = cap;
    P.cand_enc = d_cenc.as<unsigned long long>();
    P.cand_slot = d_cslot.as<uint32_t>();
    P.bslot = d_bslot.as<uint32_t>();
    P.brow = d_brow.as<int64_t>();
    P.bbox = d_bbox.as<int32_t>();
    P.hist = d_hist.as<int32_t>();
    P.order = d_order.as<int64_t>();
    P.st = d_st.as<gs::MaxminState>();
    HIP_TRY(gs::launch_maxmin_init(D, P, s));
    HIP_TRY(gs::launch_maxmin_batch(D, P, s));
    // Chunks of rounds; between chunks the state comes back.  The host
    // takes over when the last chunk appended fewer than kMinRate points per
    // round on average (each round costs a few kernel launches, the host
    // appends a point in about a microsecond).
    const char *ev = std::getenv("GS_MAXMIN_HOST_AFTER");   // test hook: hand over after this many rounds
    const int64_t host_after = ev ? std::atoll(ev) : -1;
    constexpr double kMinRate = 8.0;
    gs::MaxminState st;
    std::memset(&st, 0, sizeof(st));
    int64_t chunk = 64, prev_s = 1, prev_rounds = 0, enqueued = 1;
    for (;;) {
        int64_t run = chunk;
        if (host_after >= 0) run = std::min(run, std::max<int64_t>(0, host_after - enqueued));
        if (run == 0) {
            HIP_TRY(hipMemcpyAsync(&st, d_st.p, sizeof(st), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            break;
        }
        for (int64_t r = 0; r < run; r++) HIP_TRY(gs::launch_maxmin_round(D, P, s));
        enqueued += run;
        HIP_TRY(hipMemcpyAsync(&st, d_st.p, sizeof(st), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (st.done) break;
        const double rate = (double)(st.s - prev_s) / (double)std::max<int64_t>(1, st.rounds - prev_rounds);
        // early rounds append few points whatever the input: test the rate from 256 rounds on
        if (st.rounds >= 256 && rate < kMinRate) break;
        prev_s = st.s;
        prev_rounds = st.rounds;
        // enough rounds for the rest at the last chunk's rate (rates grow
        // along the traversal), so that few rounds are launched after the
        // end (they return at once, but each still costs its launches)
        const double need = 1.25 * (double)(n - st.s) / std::max(rate, 1.0) + 4.0;
        chunk = std::min<int64_t>(512, std::max<int64_t>(16, (int64_t)need));
    }
    if (info) { info[0] = st.rounds; info[1] = st.retries; info[2] = st.s; }
    order.resize((size_t)st.s);
    HIP_TRY(hipMemcpyAsync(order.data(), d_order.p, sizeof(int64_t) * st.s, hipMemcpyDeviceToHost, s));
    if (st.done) {
        HIP_TRY(hipStreamSynchronize(s));
        return GS_OK;
    }
    // keys around the last batch are lowered by the next round's first
    // step: run it alone so that every key is exact
    HIP_TRY(gs::launch_maxmin_update(D, P, s));
    std::vector<unsigned long long> enc((size_t)n);
    HIP_TRY(hipMemcpyAsync(enc.data(), d_enc.p, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<double> key((size_t)n);
    for (int64_t q = 0; q < n; q++) {
        const unsigned long long e = enc[(size_t)q];
        if (e == 0) {
            key[(size_t)q] = -1.0;
        } else {
            const unsigned long long kb = e - 1;
            std::memcpy(&key[(size_t)q], &kb, sizeof(double));
        }
    }
    maxmin_traverse(S, key, order);
    return GS_OK;
}

int host_threads() {
    unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 64u));
}

// Morton (Z-order) index of a cell: the bits of the D axes interleaved
// (16 bits per axis in 2-D, 10 in 3-D).
uint32_t morton_key(const uint32_t *X, int D) {
    if (D == 2) {
        auto spread = [](uint32_t v) {
            v = (v | (v << 8)) & 0x00FF00FFu;
            v = (v | (v << 4)) & 0x0F0F0F0Fu;
            v = (v | (v << 2)) & 0x33333333u;
            return (v | (v << 1)) & 0x55555555u;
        };
        return spread(X[0]) | (spread(X[1]) << 1);
    }
    auto spread3 = [](uint32_t v) {
        v = (v | (v << 16)) & 0x030000FFu;
        v = (v | (v << 8)) & 0x0300F00Fu;
        v = (v | (v << 4)) & 0x030C30C3u;
        return (v | (v << 2)) & 0x09249249u;
    };
    return spread3(X[0]) | (spread3(X[1]) << 1) | (spread3(X[2]) << 2);
}

// Points ordered along the Morton curve of their bounding box (ties by
// index): an LSD radix sort of (key << 32 | index) in 11-bit digits.
std::vector<int64_t> curve_order(int D, const double *c, int64_t n) {
    const int bits = D == 2 ? 16 : 10;
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) { mn[d] = INFINITY; mx[d] = -INFINITY; }
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < D; d++) {
            mn[d] = std::fmin(mn[d], c[i * D + d]);
            mx[d] = std::fmax(mx[d], c[i * D + d]);
        }
    const double top = (double)((1u << bits) - 1);
    std::vector<uint64_t> kv((size_t)n), tmp((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        uint32_t X[3] = {0, 0, 0};
        for (int d = 0; d < D; d++) {
            const double e = mx[d] - mn[d];
            const double f = e > 0.0 ? (c[i * D + d] - mn[d]) / e * top : 0.0;
            X[d] = (uint32_t)std::min(top, std::max(0.0, f));
        }
        kv[(size_t)i] = ((uint64_t)morton_key(X, D) << 32) | (uint64_t)i;
    }
    auto pass = [&](const std::vector<uint64_t> &src, std::vector<uint64_t> &dst, int shift) {
        std::vector<int64_t> cnt(2049, 0);
        for (int64_t i = 0; i < n; i++) cnt[((src[(size_t)i] >> shift) & 2047u) + 1]++;
        for (size_t j = 1; j < cnt.size(); j++) cnt[j] += cnt[j - 1];
        for (int64_t i = 0; i < n; i++) dst[(size_t)cnt[(src[(size_t)i] >> shift) & 2047u]++] = src[(size_t)i];
    };
    pass(kv, tmp, 32);                    // key bits 0-10, 11-21, 22-31
    pass(tmp, kv, 43);
    pass(kv, tmp, 54);
    kv.swap(tmp);
    std::vector<int64_t> perm((size_t)n);
    for (int64_t h = 0; h < n; h++) perm[(size_t)h] = (int64_t)(kv[(size_t)h] & 0xffffffffu);
    return perm;
}

}  // namespace

extern "C" {

int32_t gs_maxmin_order(int32_t dim, const double *coords, int64_t n, int64_t *order) {
    if (dim != 2 && dim != 3) return fail(GS_ERR_DIMENSION_MISMATCH, "dim must be 2 or 3");
    if (n < 0) return fail(GS_ERR_INVALID_PARAMETER, "n < 0");
    if (n == 0) return GS_OK;
    if (!coords || !order) return fail(GS_ERR_INVALID_PARAMETER, "NULL buffer");
    for (int64_t i = 0; i < n * dim; i++)
        if (!std::isfinite(coords[i])) return fail(GS_ERR_INVALID_PARAMETER, "non-finite coordinate");
    std::vector<int64_t> o;
    maxmin_order(dim, coords, n, o);
    std::copy(o.begin(), o.end(), order);
    return GS_OK;
}

int32_t gs_maxmin_order_gpu(gs_ctx *ctx, int32_t dim, const double *coords, int64_t n,
                            int64_t *order, int64_t *info) {
    if (!ctx) return fail(GS_ERR_INVALID_PARAMETER, "ctx is NULL");
    if (dim != 2 && dim != 3) return fail(GS_ERR_DIMENSION_MISMATCH, "dim must be 2 or 3");
    if (n < 0) return fail(GS_ERR_INVALID_PARAMETER, "n < 0");
    if (info) { info[0] = 0; info[1] = 0; info[2] = 0; }
    if (n == 0) return GS_OK;
    if (!coords || !order) return fail(GS_ERR_INVALID_PARAMETER, "NULL buffer");
    for (int64_t i = 0; i < n * dim; i++)
        if (!std::isfinite(coords[i])) return fail(GS_ERR_INVALID_PARAMETER, "non-finite coordinate");
    std::vector<int64_t> o;
    int32_t rc = maxmin_order_gpu(ctx, dim, coords, n, o, info, ctx->stream);
    if (rc != GS_OK) return rc;
    std::copy(o.begin(), o.end(), order);
    return GS_OK;
}

int32_t gs_vecchia_plan_new(gs_ctx *ctx, int32_t dim, const double *coords, int64_t n, int64_t m,
                            const int64_t *order, gs_vplan **out) {
    if (!out) return fail(GS_ERR_INVALID_PARAMETER, "out is NULL");
    *out = nullptr;
    if (!ctx) return fail(GS_ERR_INVALID_PARAMETER, "ctx is NULL");
    if (dim != 2 && dim != 3) return fail(GS_ERR_DIMENSION_MISMATCH, "dim must be 2 or 3");
    if (n < 2) return fail(GS_ERR_INSUFFICIENT_DATA, "Vecchia requires at least 2 points");
    if (m <= 0) return fail(GS_ERR_INVALID_PARAMETER, "conditioning size m must be at least 1");
    if (!coords) return fail(GS_ERR_INVALID_PARAMETER, "coords is NULL");
    for (int64_t i = 0; i < n * dim; i++)
        if (!std::isfinite(coords[i])) return fail(GS_ERR_INVALID_PARAMETER, "non-finite coordinate");
    int64_t pair[2];
    if (duplicate_pair(dim, coords, n, pair))
        return fail(GS_ERR_DUPLICATE_POINTS, "duplicate points " + std::to_string(pair[0]) + ", " +
                                                 std::to_string(pair[1]));
    if (m > 64) return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: conditioning size above 64");
    std::unique_ptr<gs_vplan> P(new gs_vplan());
    P->ctx = ctx;
    P->D = dim;
    P->n = n;
    P->m = m;
    if (order) {
        std::vector<char> seen((size_t)n, 0);
        for (int64_t i = 0; i < n; i++) {
            if (order[i] < 0 || order[i] >= n || seen[(size_t)order[i]])
                return fail(GS_ERR_INVALID_PARAMETER, "order is not a permutation of 0..n");
            seen[(size_t)order[i]] = 1;
        }
        P->order.assign(order, order + n);
    } else {
        HIP_TRY(hipSetDevice(ctx->device));
        int32_t rc = maxmin_order_gpu(ctx, dim, coords, n, P->order, nullptr, ctx->stream);
        if (rc != GS_OK) return rc;
    }
    // Nearest predecessors (vecchia.rs:262-281): all points inserted in
    // ordered position; query k only sees positions < k.
    std::vector<double> opts((size_t)(n * dim));
    for (int64_t k = 0; k < n; k++)
        for (int d = 0; d < dim; d++) opts[(size_t)(k * dim + d)] = coords[P->order[(size_t)k] * dim + d];
    double mn[3], mx[3];
    for (int d = 0; d < dim; d++) { mn[d] = coords[d]; mx[d] = coords[d]; }
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < dim; d++) {
            mn[d] = std::fmin(mn[d], coords[i * dim + d]);
            mx[d] = std::fmax(mx[d], coords[i * dim + d]);
        }
    BGrid G;
    G.init(dim, mn, mx, n);
    G.build_all(opts.data(), n);
    P->nb.assign((size_t)(n * m), -1);
    P->nb_count.assign((size_t)n, 0);
    // canonical predecessors on the GPU; ties at the m-th distance are
    // replayed here in the reference's visiting order
    std::vector<int32_t> gpos((size_t)(n * m)), gflag((size_t)n);
    {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        DBuf d_pts, d_start, d_items, d_out, d_cnt, d_flag;
        HIP_TRY(d_pts.ensure(sizeof(double) * n * dim));
        HIP_TRY(d_start.ensure(sizeof(int64_t) * G.start.size()));
        HIP_TRY(d_items.ensure(sizeof(int32_t) * n));
        HIP_TRY(d_out.ensure(sizeof(int32_t) * n * m));
        HIP_TRY(d_cnt.ensure(sizeof(int32_t) * n));
        HIP_TRY(d_flag.ensure(sizeof(int32_t) * n));
        HIP_TRY(hipMemcpyAsync(d_pts.p, opts.data(), sizeof(double) * n * dim, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_start.p, G.start.data(), sizeof(int64_t) * G.start.size(),
                               hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_items.p, G.items.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
        gs::PredParams pp;
        std::memset(&pp, 0, sizeof(pp));
        pp.pts = d_pts.as<double>();
        pp.n = n;
        pp.m = (int32_t)m;
        pp.brute = 1024;
        for (int d = 0; d < 3; d++) {
            pp.origin[d] = d < dim ? G.origin[d] : 0.0;
            pp.nn[d] = d < dim ? G.nn[d] : 1;
            pp.strides[d] = d < dim ? G.strides[d] : 0;
        }
        pp.cell = G.cell;
        pp.start = d_start.as<int64_t>();
        pp.items = d_items.as<int32_t>();
        pp.out = d_out.as<int32_t>();
        pp.count = d_cnt.as<int32_t>();
        pp.flag = d_flag.as<int32_t>();
        HIP_TRY(gs::launch_pred_knn(dim, pp, s));
        HIP_TRY(hipMemcpyAsync(gpos.data(), d_out.p, sizeof(int32_t) * n * m, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(P->nb_count.data(), d_cnt.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(gflag.data(), d_flag.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    std::atomic<int64_t> next(0);
    auto worker = [&]() {
        KHeap h;
        std::vector<uint32_t> res;
        for (;;) {
            int64_t s = next.fetch_add(4096);
            if (s >= n) break;
            int64_t e = std::min(n, s + 4096);
            for (int64_t k = s; k < e; k++) {
                int32_t *row = &P->nb[(size_t)(k * m)];
                if (gflag[(size_t)k]) {
                    grid_knn(G, &opts[(size_t)(k * dim)], m, k, h, res);
                    P->nb_count[(size_t)k] = (int32_t)res.size();
                    for (size_t q = 0; q < res.size(); q++) row[q] = (int32_t)P->order[res[q]];
                } else {
                    for (int32_t q = 0; q < P->nb_count[(size_t)k]; q++)
                        row[q] = (int32_t)P->order[(size_t)gpos[(size_t)(k * m + q)]];
                }
            }
        }
    };
    int nt = host_threads();
    std::vector<std::thread> th;
    for (int i = 1; i < nt; i++) th.emplace_back(worker);
    worker();
    for (auto &t : th) t.join();
    for (int64_t k = 0; k < n; k++) P->max_count = std::max(P->max_count, P->nb_count[(size_t)k]);
    {
        double e2 = 0.0;      // mle_fit_with_plan's extent (vecchia.rs:965-974)
        for (int d = 0; d < dim; d++) e2 += (mx[d] - mn[d]) * (mx[d] - mn[d]);
        P->extent = std::sqrt(e2);
    }

    // device arrays in curve order (the plan's terms are independent, so
    // the likelihood is the same sum in another order)
    P->perm = curve_order(dim, coords, n);
    {
        std::vector<int64_t> rank((size_t)n), pos((size_t)n), ident((size_t)n);
        for (int64_t k = 0; k < n; k++) rank[(size_t)P->order[(size_t)k]] = k;
        for (int64_t h = 0; h < n; h++) pos[(size_t)P->perm[(size_t)h]] = h;
        std::vector<double> hc((size_t)(n * dim));
        std::unique_ptr<int32_t[]> hnb(new int32_t[(size_t)(n * m)]);
        std::vector<int32_t> hcnt((size_t)n);
        std::atomic<int64_t> nxt(0);
        auto fill = [&]() {
            for (;;) {
                const int64_t s0 = nxt.fetch_add(8192);
                if (s0 >= n) break;
                const int64_t e0 = std::min(n, s0 + 8192);
                for (int64_t h = s0; h < e0; h++) {
                    const int64_t i = P->perm[(size_t)h];
                    const int64_t k = rank[(size_t)i];
                    ident[(size_t)h] = h;
                    for (int d = 0; d < dim; d++) hc[(size_t)(h * dim + d)] = coords[i * dim + d];
                    const int32_t c = P->nb_count[(size_t)k];
                    hcnt[(size_t)h] = c;
                    for (int32_t q = 0; q < c; q++)
                        hnb[(size_t)(h * m + q)] = (int32_t)pos[(size_t)P->nb[(size_t)(k * m + q)]];
                    for (int64_t q = c; q < m; q++) hnb[(size_t)(h * m + q)] = 0;
                }
            }
        };
        std::vector<std::thread> ft;
        for (int i = 1; i < nt; i++) ft.emplace_back(fill);
        fill();
        for (auto &t : ft) t.join();
        HIP_TRY(hipSetDevice(ctx->device));
        HIP_TRY(P->coords.ensure(sizeof(double) * n * dim));
        HIP_TRY(P->d_order.ensure(sizeof(int64_t) * n));
        HIP_TRY(P->d_nb.ensure(sizeof(int32_t) * n * m));
        HIP_TRY(P->d_count.ensure(sizeof(int32_t) * n));
        HIP_TRY(hipMemcpy(P->coords.p, hc.data(), sizeof(double) * n * dim, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->d_order.p, ident.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->d_nb.p, hnb.get(), sizeof(int32_t) * n * m, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(P->d_count.p, hcnt.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    }
    *out = P.release();
    return GS_OK;
}

void gs_vecchia_plan_free(gs_vplan *p) {
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    delete p;
}

int32_t gs_vecchia_plan_get(const gs_vplan *p, int64_t *order, int32_t *nb, int32_t *nb_count) {
    if (!p) return fail(GS_ERR_INVALID_PARAMETER, "plan is NULL");
    if (order) std::memcpy(order, p->order.data(), sizeof(int64_t) * p->n);
    if (nb) std::memcpy(nb, p->nb.data(), sizeof(int32_t) * p->n * p->m);
    if (nb_count) std::memcpy(nb_count, p->nb_count.data(), sizeof(int32_t) * p->n);
    return GS_OK;
}

}  // extern "C"

namespace {

int32_t vecchia_batch(gs_vplan *p, const double *values, const gs_model *models,
                      int32_t n_models, int32_t precision, double *out, int32_t *status,
                      hipStream_t stream) {
    if (!p) return fail(GS_ERR_INVALID_PARAMETER, "plan is NULL");
    if (n_models < 0) return fail(GS_ERR_INVALID_PARAMETER, "n_models < 0");
    if (n_models == 0) return GS_OK;
    if (!values || !models || !out) return fail(GS_ERR_INVALID_PARAMETER, "NULL buffer");
    if (precision != GS_PREC_FP64 && precision != GS_PREC_FP32_REFINED)
        return fail(GS_ERR_INVALID_PARAMETER, "unknown precision");
    const int64_t n = p->n;
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite(values[i])) return fail(GS_ERR_INVALID_PARAMETER, "non-finite value");
    for (int32_t i = 0; i < n_models; i++) {
        int32_t st = validate_model(&models[i], p->D, false);
        if (st != GS_OK) return st;
    }
    std::lock_guard<std::mutex> lk(p->mu);
    HIP_TRY(hipSetDevice(p->ctx->device));
    hipStream_t s = stream ? stream : p->ctx->stream;
    // z = v - mean (vecchia.rs:290-292), mean as a left-to-right sum.
    double mean = -0.0;
    for (int64_t i = 0; i < n; i++) mean += values[i];
    mean = mean / (double)n;
    std::vector<double> zc((size_t)n);
    for (int64_t h = 0; h < n; h++) zc[(size_t)h] = values[p->perm[(size_t)h]] - mean;   // by position
    std::vector<DevModel> dm((size_t)n_models);
    for (int32_t i = 0; i < n_models; i++) dm[(size_t)i] = resolve_model(models[i], p->D, p->ctx->nodes.as<gs::BesselNodes>());
    const int NB = pad_nb(std::max<int32_t>(p->max_count, 1));
    const int32_t ppb = 64;   // points per block
    const int64_t n_blocks = (n + ppb - 1) / ppb;
    HIP_TRY(p->zc.ensure(sizeof(double) * n));
    HIP_TRY(p->models.ensure(sizeof(DevModel) * n_models));
    HIP_TRY(p->partial.ensure(sizeof(double) * n_blocks * gs::kMaxModelsPerLaunch));
    HIP_TRY(p->out.ensure(sizeof(double) * n_models));
    HIP_TRY(p->status.ensure(sizeof(int32_t) * n_models));
    HIP_TRY(hipMemcpyAsync(p->zc.p, zc.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(p->models.p, dm.data(), sizeof(DevModel) * n_models, hipMemcpyHostToDevice, s));
    HIP_TRY(gs::launch_zero_i32(p->status.as<int32_t>(), n_models, s));
    for (int32_t b0 = 0; b0 < n_models; b0 += gs::kMaxModelsPerLaunch) {
        int32_t nm = std::min<int32_t>(gs::kMaxModelsPerLaunch, n_models - b0);
        bool lean = true;
        for (int32_t i = b0; i < b0 + nm; i++) lean = lean && gs::model_is_lean(dm[(size_t)i]);
        gs::VecchiaParams vp;
        std::memset(&vp, 0, sizeof(vp));
        vp.coords = p->coords.as<double>();
        vp.zc = p->zc.as<double>();
        vp.order = p->d_order.as<int64_t>();
        vp.nb = p->d_nb.as<int32_t>();
        vp.nb_count = p->d_count.as<int32_t>();
        vp.n = n;
        vp.m = (int32_t)p->m;
        vp.n_models = nm;
        vp.models = p->models.as<DevModel>() + b0;
        vp.partial = p->partial.as<double>();
        vp.status = p->status.as<int32_t>() + b0;
        vp.points_per_block = ppb;
        const int variant = schedule_for("vecchia", p->D, NB, 0, lean, false).variant;
        if (precision == GS_PREC_FP32_REFINED && lean) {
            // FP32-refined conditionals; a model with an untrusted one is
            // evaluated again in FP64 below
            HIP_TRY(p->refine_fail.ensure(sizeof(int32_t) * gs::kMaxModelsPerLaunch));
            HIP_TRY(gs::launch_zero_i32(p->refine_fail.as<int32_t>(), nm, s));
            vp.refine_fail = p->refine_fail.as<int32_t>();
            vp.fb_ratio = kFallbackRatioVecchia;
            hipError_t e = gs::launch_vecchia_fp32r(p->D, NB, vp, n_blocks, s);
            if (e == hipErrorInvalidValue) return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: Vecchia size");
            HIP_TRY(e);
            g_solve_kernel = "vecchia_fp32r_kernel";
            HIP_TRY(gs::launch_vecchia_reduce(p->partial.as<double>(), n_blocks, nm, p->out.as<double>() + b0, s));
            std::vector<int32_t> bad((size_t)nm);
            HIP_TRY(hipMemcpyAsync(bad.data(), p->refine_fail.p, sizeof(int32_t) * nm,
                                   hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            std::vector<int32_t> redo;
            for (int32_t i = 0; i < nm; i++)
                if (bad[(size_t)i]) redo.push_back(b0 + i);
            if (redo.empty()) continue;
            // the flagged models, packed, through the FP64 kernel
            const int32_t nr = (int32_t)redo.size();
            std::vector<DevModel> rm((size_t)nr);
            for (int32_t i = 0; i < nr; i++) rm[(size_t)i] = dm[(size_t)redo[(size_t)i]];
            HIP_TRY(p->models_redo.ensure(sizeof(DevModel) * nr));
            HIP_TRY(p->out_redo.ensure(sizeof(double) * nr));
            HIP_TRY(p->status_redo.ensure(sizeof(int32_t) * nr));
            HIP_TRY(hipMemcpyAsync(p->models_redo.p, rm.data(), sizeof(DevModel) * nr,
                                   hipMemcpyHostToDevice, s));
            HIP_TRY(gs::launch_zero_i32(p->status_redo.as<int32_t>(), nr, s));
            gs::VecchiaParams vr = vp;
            vr.n_models = nr;
            vr.models = p->models_redo.as<DevModel>();
            vr.status = p->status_redo.as<int32_t>();
            vr.refine_fail = nullptr;
            HIP_TRY(gs::launch_vecchia(p->D, NB, variant, lean, vr, n_blocks, s));
            HIP_TRY(gs::launch_vecchia_reduce(p->partial.as<double>(), n_blocks, nr,
                                              p->out_redo.as<double>(), s));
            std::vector<double> ro((size_t)nr);
            std::vector<int32_t> rs((size_t)nr);
            HIP_TRY(hipMemcpyAsync(ro.data(), p->out_redo.p, sizeof(double) * nr, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(rs.data(), p->status_redo.p, sizeof(int32_t) * nr, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            for (int32_t i = 0; i < nr; i++) {
                HIP_TRY(hipMemcpyAsync(p->out.as<double>() + redo[(size_t)i], &ro[(size_t)i],
                                       sizeof(double), hipMemcpyHostToDevice, s));
                HIP_TRY(hipMemcpyAsync(p->status.as<int32_t>() + redo[(size_t)i], &rs[(size_t)i],
                                       sizeof(int32_t), hipMemcpyHostToDevice, s));
            }
            HIP_TRY(hipStreamSynchronize(s));
            continue;
        }
        // 17-32-point conditioning sets with one lean structure of one kind
        // and anisotropy flag across the batch: factorization on the matrix
        // cores (GS_VECCHIA_MFMA=0 keeps the vector-unit kernel, for comparison)
        bool mf_aniso = false;
        const char *mf_env = std::getenv("GS_VECCHIA_MFMA");
        const int mf_kind = (NB == 32 && !(mf_env && mf_env[0] == '0'))
                                ? gs::vecchia_mfma_kind(dm.data() + b0, nm, p->max_count, &mf_aniso)
                                : -1;
        hipError_t e = mf_kind >= 0 ? gs::launch_vecchia_mfma(p->D, mf_aniso, mf_kind, vp, n_blocks, s)
                                    : gs::launch_vecchia(p->D, NB, variant, lean, vp, n_blocks, s);
        if (e == hipErrorInvalidValue) return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: Vecchia size");
        HIP_TRY(e);
        g_solve_kernel = mf_kind >= 0 ? "vecchia_mfma_kernel" : "vecchia_kernel";
        HIP_TRY(gs::launch_vecchia_reduce(p->partial.as<double>(), n_blocks, nm, p->out.as<double>() + b0, s));
    }
    std::vector<int32_t> st((size_t)n_models);
    HIP_TRY(hipMemcpyAsync(out, p->out.p, sizeof(double) * n_models, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(st.data(), p->status.p, sizeof(int32_t) * n_models, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int32_t i = 0; i < n_models; i++) {
        if (st[(size_t)i] != 0) out[i] = std::numeric_limits<double>::quiet_NaN();
        if (status) status[i] = st[(size_t)i];
    }
    return GS_OK;
}

}  // namespace

extern "C" {

int32_t gs_vecchia_loglik_batch(gs_vplan *p, const double *values, const gs_model *models,
                                int32_t n_models, int32_t precision, double *out, int32_t *status) {
    return vecchia_batch(p, values, models, n_models, precision, out, status, nullptr);
}

int32_t gs_vecchia_loglik_grad_batch(gs_vplan *p, const double *values, const gs_model *models,
                                     int32_t n_models, int32_t precision, double *loglik,
                                     double *grad, int32_t *status) {
    if (!p) return fail(GS_ERR_INVALID_PARAMETER, "plan is NULL");
    if (n_models < 0) return fail(GS_ERR_INVALID_PARAMETER, "n_models < 0");
    if (n_models == 0) return GS_OK;
    if (!values || !models || !loglik || !grad) return fail(GS_ERR_INVALID_PARAMETER, "NULL buffer");
    if (precision != GS_PREC_FP64 && precision != GS_PREC_FP32_REFINED)
        return fail(GS_ERR_INVALID_PARAMETER, "unknown precision");
    // the gradient runs in FP64 for either precision
    if (p->max_count > 32)
        return fail(GS_ERR_INVALID_PARAMETER,
                    "unsupported on GPU: gradient with conditioning sets above 32 points");
    const int64_t n = p->n;
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite(values[i])) return fail(GS_ERR_INVALID_PARAMETER, "non-finite value");
    // models the reference's loglik_grad_with_plan answers (g_and_dg kinds,
    // one isotropic structure); the others get status InvalidParameter
    std::vector<int32_t> st((size_t)n_models, 0), elig;
    for (int32_t i = 0; i < n_models; i++) {
        int32_t v = validate_model(&models[i], p->D, false);
        if (v != GS_OK) return v;
        const gs_model &m = models[i];
        const int32_t k = m.n_structures == 1 ? m.structures[0].kind : -1;
        const bool ok = m.n_structures == 1 && !m.structures[0].has_anis &&
                        (k == GS_SPHERICAL || k == GS_EXPONENTIAL || k == GS_GAUSSIAN ||
                         k == GS_MATERN15 || k == GS_MATERN25);
        if (ok) elig.push_back(i);
        else st[(size_t)i] = GS_ERR_INVALID_PARAMETER;
    }
    std::vector<double> res((size_t)n_models * 4, std::numeric_limits<double>::quiet_NaN());
    if (!elig.empty()) {
        std::lock_guard<std::mutex> lk(p->mu);
        HIP_TRY(hipSetDevice(p->ctx->device));
        hipStream_t s = p->ctx->stream;
        double mean = -0.0;        // z = v - mean (vecchia.rs:437-439)
        for (int64_t i = 0; i < n; i++) mean += values[i];
        mean = mean / (double)n;
        std::vector<double> zc((size_t)n);
        for (int64_t h = 0; h < n; h++) zc[(size_t)h] = values[p->perm[(size_t)h]] - mean;   // by position
        const int32_t ne = (int32_t)elig.size();
        std::vector<DevModel> dm((size_t)ne);
        for (int32_t i = 0; i < ne; i++)
            dm[(size_t)i] = resolve_model(models[elig[(size_t)i]], p->D, p->ctx->nodes.as<gs::BesselNodes>());
        const int NB = pad_nb(std::max<int32_t>(p->max_count, 1));
        const int32_t ppb = 64;
        const int64_t n_blocks = (n + ppb - 1) / ppb;
        HIP_TRY(p->zc.ensure(sizeof(double) * n));
        HIP_TRY(p->models.ensure(sizeof(DevModel) * ne));
        HIP_TRY(p->partial.ensure(sizeof(double) * n_blocks * gs::kMaxModelsPerLaunch * 4));
        HIP_TRY(p->out.ensure(sizeof(double) * ne * 4));
        HIP_TRY(p->status.ensure(sizeof(int32_t) * ne));
        HIP_TRY(hipMemcpyAsync(p->zc.p, zc.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(p->models.p, dm.data(), sizeof(DevModel) * ne, hipMemcpyHostToDevice, s));
        HIP_TRY(gs::launch_zero_i32(p->status.as<int32_t>(), ne, s));
        for (int32_t b0 = 0; b0 < ne; b0 += gs::kMaxModelsPerLaunch) {
            const int32_t nm = std::min<int32_t>(gs::kMaxModelsPerLaunch, ne - b0);
            gs::VecchiaParams vp;
            std::memset(&vp, 0, sizeof(vp));
            vp.coords = p->coords.as<double>();
            vp.zc = p->zc.as<double>();
            vp.order = p->d_order.as<int64_t>();
            vp.nb = p->d_nb.as<int32_t>();
            vp.nb_count = p->d_count.as<int32_t>();
            vp.n = n;
            vp.m = (int32_t)p->m;
            vp.n_models = nm;
            vp.models = p->models.as<DevModel>() + b0;
            vp.partial = p->partial.as<double>();
            vp.status = p->status.as<int32_t>() + b0;
            vp.points_per_block = ppb;
            // one kind across the batch, 17-32 conditioning slots: the
            // matrix-core kernel (GS_VECCHIA_MFMA=0 keeps the vector one)
            const char *mf_env = std::getenv("GS_VECCHIA_MFMA");
            bool mf_aniso = false;
            const int mf_kind = (NB == 32 && !(mf_env && mf_env[0] == '0'))
                ? gs::vecchia_mfma_kind(dm.data() + b0, nm, p->max_count, &mf_aniso) : -1;
            hipError_t e = mf_kind >= 0 && !mf_aniso ? gs::launch_vecchia_grad_mfma(p->D, mf_kind, vp, n_blocks, s)
                                                     : gs::launch_vecchia_grad(p->D, NB, vp, n_blocks, s);
            if (e == hipErrorInvalidValue) return fail(GS_ERR_INVALID_PARAMETER, "unsupported on GPU: Vecchia size");
            HIP_TRY(e);
            g_solve_kernel = mf_kind >= 0 && !mf_aniso ? "vecchia_grad_mfma_kernel" : "vecchia_grad_kernel";
            HIP_TRY(gs::launch_vecchia_reduce(p->partial.as<double>(), n_blocks, nm * 4,
                                              p->out.as<double>() + (int64_t)b0 * 4, s));
        }
        std::vector<double> o4((size_t)ne * 4);
        std::vector<int32_t> st_e((size_t)ne);
        HIP_TRY(hipMemcpyAsync(o4.data(), p->out.p, sizeof(double) * ne * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(st_e.data(), p->status.p, sizeof(int32_t) * ne, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int32_t i = 0; i < ne; i++) {
            const int32_t mi = elig[(size_t)i];
            st[(size_t)mi] = st_e[(size_t)i];
            if (st_e[(size_t)i] == 0)
                for (int q = 0; q < 4; q++) res[(size_t)mi * 4 + q] = o4[(size_t)i * 4 + q];
        }
    }
    for (int32_t i = 0; i < n_models; i++) {
        loglik[i] = res[(size_t)i * 4];
        for (int q = 0; q < 3; q++) grad[(size_t)i * 3 + q] = res[(size_t)i * 4 + 1 + q];
        if (status) status[i] = st[(size_t)i];
    }
    return GS_OK;
}

int32_t gs_vecchia_loglik(gs_ctx *ctx, int32_t dim, const double *coords, const double *values,
                          int64_t n, const gs_model *model, int64_t m, const int64_t *order,
                          double *out) {
    if (!out) return fail(GS_ERR_INVALID_PARAMETER, "out is NULL");
    // reject_power_model / reject_invalid_dim_model come first (vecchia.rs:880-881)
    int32_t st = validate_model(model, dim, false);
    if (st != GS_OK) return st;
    gs_vplan *p = nullptr;
    st = gs_vecchia_plan_new(ctx, dim, coords, n, m, order, &p);
    if (st != GS_OK) return st;
    int32_t s1 = 0;
    st = gs_vecchia_loglik_batch(p, values, model, 1, GS_PREC_FP64, out, &s1);
    gs_vecchia_plan_free(p);
    if (st != GS_OK) return st;
    if (s1 != 0) return fail(s1, "singular conditioning system");
    return GS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Request loop: kriging and Vecchia-likelihood requests over resident
// predictors and plans.  Each worker slot owns a HIP stream and a kriging
// workspace; a request is admitted, first come first served, when a slot is
// free and the device bytes it needs fit in the budget beside the requests
// already running (a request larger than the whole budget runs alone).
// Requests on different slots run concurrently, so their kernels overlap on
// the device.  Every kernel on the path is deterministic and independent of
// its launch grid, so a request's result is the same alone and in the mix.
// ---------------------------------------------------------------------------
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>

namespace {

using Clock = std::chrono::steady_clock;

struct SReq {
    int kind = 0;   // 0 kriging, 1 Vecchia likelihood
    int32_t precision = GS_PREC_FP64;
    gs_kriging *k = nullptr;
    const double *targets = nullptr;
    int64_t T = 0;
    double *value = nullptr, *variance = nullptr, *lagrange = nullptr;
    int32_t *status = nullptr;
    gs_vplan *p = nullptr;
    const double *values = nullptr;
    const gs_model *models = nullptr;
    int32_t n_models = 0;
    double *out = nullptr;
    int32_t *vstatus = nullptr;
    int64_t bytes = 0;
    Clock::time_point t_submit, t_start, t_end;
    int32_t result = GS_OK;
    std::string error;
    bool done = false;
};

// Device bytes a kriging request needs beyond the resident data: the
// workspace and staging of predict_host (neighbour lists, search order and
// sort scratch, outputs) or the global scratch.
int64_t krige_bytes(const gs_kriging *k, int64_t T) {
    if (k->global) {
        const int64_t N = k->n + k->M;
        int64_t tc = ((int64_t)1 << 30) / (2 * N * (int64_t)sizeof(double));
        tc = std::max<int64_t>(64, tc / 64 * 64);
        tc = std::min<int64_t>(tc, (T + 63) / 64 * 64);
        return T * (k->D * 8 + 36) + 2 * N * tc * 8;
    }
    const int64_t stride = std::max<int64_t>(k->stride, 1);
    return T * (k->D * 8 + 36 + stride * 4 + 4 + 16 + 8 + 4 + 16 + (k->per_oct > 0 ? 12 : 0)) +
           (int64_t)gs::tile_sort_workspace(T, ncells_of(k));
}

int64_t vecchia_bytes(const gs_vplan *p, int32_t n_models) {
    const int64_t n_blocks = (p->n + 63) / 64;
    return p->n * 8 + (int64_t)n_models * (int64_t)(sizeof(DevModel) + 12) +
           n_blocks * gs::kMaxModelsPerLaunch * 8;
}

}  // namespace

struct gs_server {
    gs_ctx *ctx = nullptr;
    int64_t budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<SReq>> queue;
    std::map<int64_t, std::shared_ptr<SReq>> tickets;
    int64_t next_ticket = 1;
    int64_t in_flight_bytes = 0;
    int running = 0;
    bool stopping = false;
    struct Slot {
        hipStream_t stream = nullptr;
        KWork w;
    };
    std::vector<std::unique_ptr<Slot>> slots;
    std::vector<std::thread> threads;

    void worker(Slot *slot) {
        (void)hipSetDevice(ctx->device);
        for (;;) {
            std::shared_ptr<SReq> r;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    if (stopping && queue.empty()) return true;
                    if (queue.empty()) return false;
                    const int64_t b = queue.front()->bytes;
                    return running == 0 || in_flight_bytes + b <= budget;
                });
                if (queue.empty()) return;   // stopping
                r = queue.front();
                queue.pop_front();
                in_flight_bytes += r->bytes;
                running++;
                r->t_start = Clock::now();
            }
            int32_t st;
            if (r->kind == 0)
                st = predict_host(r->k, slot->w, r->targets, r->T, r->precision, r->value,
                                  r->variance, r->lagrange, r->status, nullptr, Extra(),
                                  slot->stream, false);
            else
                st = vecchia_batch(r->p, r->values, r->models, r->n_models, r->precision, r->out,
                                   r->vstatus, slot->stream);
            {
                std::lock_guard<std::mutex> lk(mu);
                r->t_end = Clock::now();
                r->result = st;
                if (st != GS_OK) r->error = g_err;
                r->done = true;
                in_flight_bytes -= r->bytes;
                running--;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

int32_t gs_server_create(gs_ctx *ctx, int32_t n_slots, int64_t mem_budget, gs_server **out) {
    if (!out) return fail(GS_ERR_INVALID_PARAMETER, "out is NULL");
    *out = nullptr;
    if (!ctx) return fail(GS_ERR_INVALID_PARAMETER, "ctx is NULL");
    if (n_slots < 1 || n_slots > 16) return fail(GS_ERR_INVALID_PARAMETER, "n_slots must be 1..16");
    if (mem_budget <= 0) return fail(GS_ERR_INVALID_PARAMETER, "mem_budget must be positive");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<gs_server> S(new gs_server());
    S->ctx = ctx;
    S->budget = mem_budget;
    for (int i = 0; i < n_slots; i++) {
        std::unique_ptr<gs_server::Slot> sl(new gs_server::Slot());
        const hipError_t e = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (auto &made : S->slots) (void)hipStreamDestroy(made->stream);
            return fail(GS_ERR_DEVICE, std::string("stream creation: ") + hipGetErrorString(e));
        }
        S->slots.push_back(std::move(sl));
    }
    for (auto &sl : S->slots) {
        gs_server::Slot *raw = sl.get();
        gs_server *srv = S.get();
        S->threads.emplace_back([srv, raw] { srv->worker(raw); });
    }
    *out = S.release();
    return GS_OK;
}

void gs_server_destroy(gs_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->threads) t.join();
    (void)hipSetDevice(s->ctx->device);
    for (auto &sl : s->slots)
        if (sl->stream) (void)hipStreamDestroy(sl->stream);
    delete s;
}

int32_t gs_server_submit_krige(gs_server *s, gs_kriging *k, const double *targets, int64_t T,
                               int32_t precision, double *value, double *variance,
                               double *lagrange, int32_t *status, int64_t *ticket) {
    if (!s || !k || !ticket) return fail(GS_ERR_INVALID_PARAMETER, "NULL argument");
    if (precision != GS_PREC_FP64 && precision != GS_PREC_FP32_REFINED)
        return fail(GS_ERR_INVALID_PARAMETER, "unknown precision");
    if (T <= 0 || !targets || !value || !variance)
        return fail(GS_ERR_INVALID_PARAMETER, "empty request or NULL buffer");
    if (k->ctx != s->ctx) return fail(GS_ERR_INVALID_PARAMETER, "predictor belongs to another context");
    if (k->nv > 0)
        return fail(GS_ERR_INVALID_PARAMETER, "unsupported in the request loop: external drift");
    auto r = std::make_shared<SReq>();
    r->kind = 0;
    r->precision = precision;
    r->k = k;
    r->targets = targets;
    r->T = T;
    r->value = value;
    r->variance = variance;
    r->lagrange = lagrange;
    r->status = status;
    r->bytes = krige_bytes(k, T);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) return fail(GS_ERR_INVALID_PARAMETER, "server is stopping");
        *ticket = s->next_ticket++;
        s->tickets[*ticket] = r;
        s->queue.push_back(r);
    }
    s->cv_work.notify_all();
    return GS_OK;
}

int32_t gs_server_submit_vecchia(gs_server *s, gs_vplan *p, const double *values,
                                 const gs_model *models, int32_t n_models, int32_t precision,
                                 double *out, int32_t *status, int64_t *ticket) {
    if (!s || !p || !ticket) return fail(GS_ERR_INVALID_PARAMETER, "NULL argument");
    if (precision != GS_PREC_FP64 && precision != GS_PREC_FP32_REFINED)
        return fail(GS_ERR_INVALID_PARAMETER, "unknown precision");
    if (n_models <= 0 || !values || !models || !out)
        return fail(GS_ERR_INVALID_PARAMETER, "empty request or NULL buffer");
    if (p->ctx != s->ctx) return fail(GS_ERR_INVALID_PARAMETER, "plan belongs to another context");
    auto r = std::make_shared<SReq>();
    r->kind = 1;
    r->precision = precision;
    r->p = p;
    r->values = values;
    r->models = models;
    r->n_models = n_models;
    r->out = out;
    r->vstatus = status;
    r->bytes = vecchia_bytes(p, n_models);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) return fail(GS_ERR_INVALID_PARAMETER, "server is stopping");
        *ticket = s->next_ticket++;
        s->tickets[*ticket] = r;
        s->queue.push_back(r);
    }
    s->cv_work.notify_all();
    return GS_OK;
}

int32_t gs_server_wait(gs_server *s, int64_t ticket, double *queue_ms, double *service_ms) {
    if (!s) return fail(GS_ERR_INVALID_PARAMETER, "server is NULL");
    std::shared_ptr<SReq> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->tickets.find(ticket);
        if (it == s->tickets.end()) return fail(GS_ERR_INVALID_PARAMETER, "unknown ticket");
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->done; });
        s->tickets.erase(ticket);
    }
    if (queue_ms)
        *queue_ms = std::chrono::duration<double, std::milli>(r->t_start - r->t_submit).count();
    if (service_ms)
        *service_ms = std::chrono::duration<double, std::milli>(r->t_end - r->t_start).count();
    if (r->result != GS_OK) return fail(r->result, r->error);
    return GS_OK;
}

}  // extern "C"
