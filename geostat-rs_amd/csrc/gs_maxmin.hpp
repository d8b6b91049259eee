// Batched maxmin ordering on the GPU (gs_maxmin.hip), shared with the host
// driver in gs_host.cpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gs {

// Sizes of the candidate sort and of the key histogram.
constexpr int kMaxminCap = 2048;    // candidates sorted per round
constexpr int kMaxminBins = 1024;   // 1/256-octave key bins below the threshold

struct MaxminState {                // device-resident, rewritten every round
    unsigned long long tau_key;     // threshold (key bits, index): candidates are
    uint32_t tau_idx;               //   (key, index) >= (tau_key, tau_idx)
    uint32_t count;                 // candidates collected this round (may exceed cap)
    int64_t s;                      // points ordered so far
    int64_t rows;                   // bucket rows of the last batch's update boxes
    int32_t batch;                  // points appended by the last round
    int32_t done;
    int64_t rounds, retries;        // rounds run; rounds that raised the threshold instead
    int64_t subrounds;              // sub-rounds of the batch step, summed
};

struct MaxminParams {
    int64_t n;
    const double *cc;               // [n][D] coordinates in bucket (slot) order
    const uint32_t *id;             // [n] slot -> original index
    unsigned long long *enc;        // [n] bits(key) + 1 while unordered, 0 once ordered
    double first[3];                // the first point (nearest the centroid)
    uint32_t first_slot;
    double origin[3];
    double cell;
    int64_t nn[3], strides[3];
    const int64_t *start;           // [buckets + 1] CSR bucket offsets
    int32_t cap;                    // candidate capacity, a power of two <= kMaxminCap
    unsigned long long *cand_enc;   // [cap]
    uint32_t *cand_slot;            // [cap]
    uint32_t *bslot;                // [cap] slots of the last batch
    int64_t *brow;                  // [cap + 1] exclusive scan of the batch's update rows
    int32_t *bbox;                  // [cap][6] bucket box per batch point (lo xyz, hi xyz)
    int32_t *hist;                  // [kMaxminBins], zero between rounds
    int64_t *order;                 // [n] original indices in maxmin order
    MaxminState *st;
};

// keys to the first point + the first histogram (hist must be zero)
hipError_t launch_maxmin_init(int D, const MaxminParams &p, hipStream_t s);
// the batch step alone (after init: picks the first threshold)
hipError_t launch_maxmin_batch(int D, const MaxminParams &p, hipStream_t s);
// one round: update around the last batch, collect, batch
hipError_t launch_maxmin_round(int D, const MaxminParams &p, hipStream_t s);
// the update step alone (makes every key exact after a round)
hipError_t launch_maxmin_update(int D, const MaxminParams &p, hipStream_t s);

}  // namespace gs
