// Device-resident fuzzy simplicial set: the output of construct_umap_graph_gpu
// (src/lib.rs:3376-3441) kept on the device, plus the CSR/edge layout the
// epoch kernels consume (UmapCsrGraph, src/training/umap_optimiser_gpu.rs:93-210).
#pragma once

#include "mrs_internal.hpp"

struct mrs_graph {
    mrs_ctx* ctx = nullptr;
    int64_t n = 0;
    int32_t k = 0;
    int32_t n_epochs = 0;
    mrs::DevBuf<float> sigma, rho;  // [n]

    // fuzzy union, one entry per (row, col) in out(row) U in(row), rows sorted by
    // col; the pattern is symmetric.  w = combine(w_ij, w_ji).
    int64_t n_union = 0;
    mrs::DevBuf<uint32_t> u_off;  // [n+1]
    mrs::DevBuf<uint32_t> u_row, u_col;
    mrs::DevBuf<float> u_w;
    mrs::DevBuf<uint32_t> u_keep;  // w > 0 && w >= max_w / n_epochs

    // filtered COO (CoordinateList after filter_weak_edges): row-major, cols ascending
    int64_t nnz = 0;
    mrs::DevBuf<uint32_t> coo_row, coo_col;
    mrs::DevBuf<float> coo_val;

    // epoch layout: every edge (i < j, row-major order) under both endpoints
    int64_t n_edges = 0;
    mrs::DevBuf<uint32_t> off;           // [n+1]
    mrs::DevBuf<uint32_t> csr_edge_idx;  // [2E]
    mrs::DevBuf<uint32_t> csr_other;     // [2E]
    mrs::DevBuf<uint4> edges;            // [E] {i, j, CSR slot under i, CSR slot under j}
    mrs::DevBuf<float> edge_w;           // [E]
    mrs::DevBuf<float> per;              // [E]
};

namespace mrs {

// Build the graph from kNN results already on the device ([n*k], self excluded).
mrs_graph* build_graph(mrs_ctx* ctx, const int32_t* knn_idx, const float* knn_dist, int64_t n,
                       int32_t k, float bandwidth, float local_connectivity, float mix_weight,
                       int32_t n_epochs);

// DensState::new on the device (src/utils/density.rs:181-201): r [n] z-scored
// original log radii and mu_sum [n] written to device buffers; returns mu_tot.
// Throws Error(MRS_DEGENERATE_LOCAL_RADII) like the reference.
float dens_state_on_device(mrs_graph* g, const int32_t* knn_idx, const float* knn_dist,
                           bool distances_squared, float* r_dev, float* mu_sum_dev);

}  // namespace mrs
