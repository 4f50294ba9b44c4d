// Device-resident optimiser state of the "adam_gpu" path: the MI355X
// counterpart of UmapGpuState (src/training/umap_optimiser_gpu.rs:278-348).
#pragma once

#include "epoch_common.hpp"
#include "mrs_internal.hpp"

struct mrs_umap_state {
    mrs_ctx* ctx = nullptr;
    int64_t n = 0;
    int32_t dim = 0;
    int64_t n_edges = 0;
    mrs_umap_optim_params params{};

    // [n*dim] row-major
    mrs::DevBuf<float> embd, m, v, grad;
    // [n] per-node "any edge fired this epoch"
    mrs::DevBuf<uint32_t> has_update;
    // CSR: every undirected edge under both endpoints, in edge order
    mrs::DevBuf<uint32_t> off;           // [n+1]
    mrs::DevBuf<uint32_t> csr_edge_idx;  // [2E]
    mrs::DevBuf<uint32_t> csr_other;     // [2E]
    // edge stream for the edge-parallel kernel, edges sorted by (i, j), i < j:
    // one 16-byte record per edge, {i, j, CSR slot of the edge under i, under j}
    mrs::DevBuf<uint4> edges;        // [E]
    mrs::DevBuf<float> per, next, mu;  // [E]
    // firing bitmask of a window of epochs (k_fire_bits): up to kFireWindow
    // planes of ceil(E / 64) words; allocated by the first run of the
    // edge-parallel loop long enough to use it
    mrs::DevBuf<uint64_t> fire_bits;
    // compact edge records of the bitmask epoch kernel (CompactRecords,
    // epoch_device.hpp), built from `edges` with the first bitmask
    mrs::DevBuf<uint2> rec_pair, run_info;
    mrs::DevBuf<uint64_t> run_start;
    mrs::DevBuf<uint32_t> word_run;

    // grad holds the edge strategy's cleared state: -0.0 everywhere, "no update
    // yet" (epoch_device.hpp); has_update is kept by the node-exact strategy
    bool grad_clean = false;

    // densMAP
    bool dens_enabled = false;
    mrs_dens_params dens_params{};
    float mu_tot = 0.0f;
    mrs::DevBuf<float> dens_r, dens_re, dens_phi, dens_w;  // [n] (or [1] when off)
    mrs::DevBuf<float> dens_pack;  // [n][4]: w, 1/phi_sum, 1/exp(re), 0 (edge kernel)
    // locality order of the radii kernel: node order[q] at position q, the CSR
    // renumbered into it (rows keep their entry order) and the embedding
    // gathered into it each density epoch; empty when not built
    mrs::DevBuf<uint32_t> dens_order, dens_off, dens_cols;
    mrs::DevBuf<float> dens_y;
    mrs::DevBuf<double> dens_partials;                     // [3 * kDensBlocks]
    mrs::DevBuf<float> dens_scalars;                       // mean, std, cov, coeff

    mrs::DevBuf<float> norm_partial;  // [n]

    // Optional per-kernel timing with HIP events on the context stream.
    // kind: 0 edge epoch, 1 adam+clear, 2 node grad, 3 adam elem, 4 schedule, 5 densMAP update,
    // 6 firing bitmask window
    bool profiling = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_events;
    double prof_ms[8] = {0};
    int64_t prof_count[8] = {0};
    // The epoch loop captured as a HIP graph, replayed with one launch.  A
    // loop is captured the second time the same (first, count, seed,
    // strategy) is run on this state: capturing costs as much host time as
    // launching directly, and the device waits through it, so a loop run
    // once (one fit) is launched directly.
    bool use_graphs = true;
    hipGraphExec_t graph_exec = nullptr;
    int64_t graph_key[4] = {-1, -1, -1, -1};
    int64_t last_key[4] = {-1, -1, -1, -1};  // the last loop launched directly
    ~mrs_umap_state();
};

namespace mrs {

constexpr int kDensBlocks = 256;
// Epochs per firing-bitmask window, and the shortest edge-parallel run that
// builds one (shorter runs read the schedule cursors in the epoch kernel).
constexpr int kFireWindow = 128;
constexpr int kFireMinEpochs = 8;

// Host CSR of an adjacency list (UmapCsrGraph::from_graph,
// src/training/umap_optimiser_gpu.rs:132-209), plus the per-edge endpoint
// arrays and local CSR slots used by the edge-parallel kernel.
struct HostCsr {
    int64_t n = 0, n_edges = 0;
    std::vector<uint32_t> off, csr_edge_idx, csr_other;
    std::vector<uint4> edges;
    std::vector<float> per, mu;
};
// Throws Error(MRS_NO_DATA / MRS_NO_GRAPH_EDGES) like the reference.
HostCsr host_csr_from_graph(int64_t n, const int64_t* adj_off, const int32_t* adj_cols,
                            const float* adj_vals);

// Allocate and upload a state from a host CSR.
mrs_umap_state* make_state(mrs_ctx* ctx, const float* embd, int64_t n, int32_t dim,
                           const HostCsr& csr, const mrs_umap_optim_params& p,
                           const mrs_dens_state* dens);

// Allocate the per-node buffers of a state whose graph arrays are already on
// the device (filled by the caller).  embd is device or host memory [n*dim].
void state_alloc_nodes(mrs_umap_state* st, const float* embd);
void state_dens_order(mrs_umap_state* st);
void dens_pack_refresh(mrs_umap_state* st);
void state_alloc_dens(mrs_umap_state* st, const mrs_dens_params& dp, const float* r_dev_or_host,
                      float mu_tot);

AdamScalars adam_scalars(const mrs_umap_optim_params& p, int32_t epoch);
bool dens_is_active(const mrs_dens_params& dp, int32_t epoch, int32_t n_epochs);

// Per-epoch launchers (asynchronous on st->ctx->stream).
void launch_grad_accum_node(mrs_umap_state* st, int32_t epoch, uint32_t seed, bool dens);
void launch_adam_elem(mrs_umap_state* st, int32_t epoch);
void launch_schedule(mrs_umap_state* st, int32_t epoch);
// radii + FP64 stats + weights, coeff on device; reference_order: the radii
// summed per node in the reference's sequential order (else 16 lanes per node)
void launch_dens_update_dev(mrs_umap_state* st, bool reference_order);
// span: edges per wavefront (64, 128 or 256); 0 = the recorded or built-in
// choice for this state (autotune.hpp).  plane: the epoch's plane of the
// firing bitmask (launch_fire_bits), or null to read and advance the schedule
// cursors in the kernel.
void launch_epoch_edges(mrs_umap_state* st, int32_t epoch, uint32_t seed, bool dens,
                        int span = 0, const uint64_t* plane = nullptr);
// Firing bitmask of epochs [first, first + count), count <= kFireWindow, into
// st->fire_bits; advances the schedule cursors past the window.
void launch_fire_bits(mrs_umap_state* st, int32_t first, int32_t count);
// The compact edge records (rec_pair, run_start, word_run, run_info) from
// st->edges; inside a pooled StreamScope.
void build_compact_records(mrs_umap_state* st);
// The span launch_epoch_edges uses for this state when given 0.
int epoch_span_choice(const mrs_umap_state* st, bool dens);
// grad to -0.0 (no update yet), has_update from / to the markers
void clear_update_marks(mrs_umap_state* st);
void derive_update_flags(mrs_umap_state* st);
void marks_from_flags(mrs_umap_state* st);
void launch_adam_clear(mrs_umap_state* st, int32_t epoch, float* norm_partial);
void run_epochs(mrs_umap_state* st, int32_t first, int32_t count, uint32_t seed, int strategy);
// Wait for and accumulate the events recorded while profiling.
void prof_collect(mrs_umap_state* st);

}  // namespace mrs
