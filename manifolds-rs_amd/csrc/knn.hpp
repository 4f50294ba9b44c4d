// Exhaustive kNN on the device (MFMA shortlist + exact FP32 re-rank).
#pragma once

#include "mrs_internal.hpp"

namespace mrs {

struct KnnTimings {
    float pack_ms = 0, shortlist_ms = 0, rerank_ms = 0;
    int splits = 0, shortlist = 0;
    double swept = 1.0;  // fraction of (query block, stage) pairs swept
    int cells = 0;       // cells of the pruned sweep (0: full sweep)
    uint32_t list_max = 0;   // longest pass-2 stage list of a query block
    double list_mean = 0.0;  // mean pass-2 stage list
    uint32_t aux_items = 0;  // pass-2 work items that fill aux lists (k_merge_aux folds them in)
};

// Caller-provided device memory for the full sweep's buffers (packed operand,
// norms, candidate lists and queues), at most max_splits database splits.
// With it the search allocates nothing, creates no event and does not wait
// for the device: it is queued on the context's stream and returns.
struct KnnScratch {
    unsigned char* base;
    size_t bytes;
    int max_splits;
};

// x: device [n][d] f32; out_idx/out_dist: device [n][k].
void knn_exhaustive_dev(mrs_ctx* ctx, const float* x, int64_t n, int32_t d, int32_t k,
                        int32_t precision, int32_t shortlist, int32_t* out_idx, float* out_dist,
                        KnnTimings* timings, const KnnScratch* scratch = nullptr);

}  // namespace mrs
