// Spectral initial layout on the device (spectral_layout, src/data/init.rs:986-1004).
#pragma once

#include "graph.hpp"

namespace mrs {

struct FitTimings {
    double knn_ms = 0, graph_ms = 0, dens_ms = 0, init_ms = 0, epochs_ms = 0, total_ms = 0;
    double knn_shortlist_ms = 0;
    int64_t n_edges = 0;
};

// Writes the initial embedding [n][dim] (device) for the filtered graph g.
// data (device [n][d] f32, may be null) places component centroids on the
// disconnected path, as the reference's optional feature matrix does.
void spectral_layout_dev(mrs_ctx* ctx, const mrs_graph* g, int32_t dim, uint64_t seed,
                         float range, float* out, const float* data = nullptr, int32_t d = 0);

void random_layout_dev(mrs_ctx* ctx, float* out, int64_t n, int32_t dim, uint64_t seed,
                       float range);

}  // namespace mrs
