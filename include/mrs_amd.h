/*
 * mrs_amd.h -- C ABI of the MI355X-native UMAP/densMAP fitting path.
 *
 * Drop-in boundary for the `umap_gpu` / `densmap_gpu` fit of
 * GregorLueg/manifolds-rs v0.3.15 with the "adam_gpu" optimiser and the
 * "exhaustive_gpu" kNN.  Each entry point names the reference interface it
 * replaces (path:line into the reference crate).  Plain pointers and sizes
 * only; float32 data; row-major buffers.  Any pointer argument may be host or
 * device memory: device pointers (e.g. from hipMalloc) are used in place,
 * host pointers are staged through the context's stream.
 *
 * Status codes (int32) map 1:1 onto ManifoldsError (src/errors.rs):
 *   MRS_OK, NoData, NoGraphEdges, DegenerateLocalRadii, CubeclUtilsError
 *   (device/allocation failures), and invalid arguments.
 * mrs_last_error() returns the message of the last failure on this thread.
 *
 * Threading: calls are synchronous on the caller's thread (as the reference's
 * are: every readback is a sync).  Use one mrs_ctx per host thread.
 */
#ifndef MRS_AMD_H
#define MRS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRS_OK 0
#define MRS_NO_DATA 1                /* ManifoldsError::NoData, src/errors.rs:78 */
#define MRS_NO_GRAPH_EDGES 2         /* ManifoldsError::NoGraphEdges, src/errors.rs:82 */
#define MRS_DEGENERATE_LOCAL_RADII 3 /* ManifoldsError::DegenerateLocalRadii, src/errors.rs:101 */
#define MRS_DEVICE_ERROR 4           /* ManifoldsError::CubeclUtilsError, src/errors.rs:54 */
#define MRS_INVALID_ARGUMENT 5

/* epoch-kernel strategies for the "adam_gpu" optimiser */
#define MRS_EPOCH_EDGE_ATOMIC 0   /* edge-parallel, FP32 atomic scatter-add (default, fastest) */
#define MRS_EPOCH_NODE_EXACT 1    /* node-parallel CSR walk, deterministic, reference order */

/* exhaustive kNN precisions */
#define MRS_KNN_FP32 0            /* FP32 MFMA candidates + exact FP32 re-rank */
#define MRS_KNN_BF16_RERANK 1     /* BF16 MFMA candidates + exact FP32 re-rank */

typedef struct mrs_ctx mrs_ctx;
typedef struct mrs_umap_state mrs_umap_state;
typedef struct mrs_graph mrs_graph;

/* UmapOptimParams<f32>, src/training/umap_optimisers.rs:22-44 */
typedef struct {
    float a, b, lr, gamma;
    int32_t n_epochs, neg_sample_rate;
    float min_dist, beta1, beta2, eps;
} mrs_umap_optim_params;

/* UmapGraphParams<f32>, src/data/graph.rs:30-61 (defaults 1e-5, 1.0, 1.0) */
typedef struct {
    float bandwidth, local_connectivity, mix_weight;
} mrs_umap_graph_params;

/* DensParams<f32>, src/utils/density.rs:69-79 */
typedef struct {
    float lambda, frac, var_shift;
} mrs_dens_params;

/* DensState<f32>, src/utils/density.rs:140-160 (r and mu_sum are [n]) */
typedef struct {
    mrs_dens_params params;
    const float *r;
    const float *mu_sum;
    float mu_tot;
} mrs_dens_state;

/* UmapParamsGpu<f32> (src/lib.rs:3218-3241) with NearestNeighbourParamsGpu's
 * dist_metric (src/data/nearest_neighbours_gpu.rs:34-62) and this build's kNN
 * and epoch knobs.  Fill with mrs_umap_params_gpu_default(). */
typedef struct {
    int32_t n_dim;
    int32_t k;
    const char *optimiser;       /* "adam_gpu" (default; unknown strings fall back to it) */
    const char *ann_type;        /* "exhaustive_gpu" (approximate types run the exact search) */
    const char *initialisation;  /* "spectral" (default) | "random" */
    float init_range;            /* NaN = None (the reference's Option<T>) */
    const char *dist_metric;     /* "euclidean" */
    int32_t knn_precision;       /* MRS_KNN_FP32 | MRS_KNN_BF16_RERANK */
    int32_t knn_shortlist;       /* re-rank shortlist (0 = 64, or 128 when k > 64) */
    mrs_umap_graph_params umap_graph_params;
    mrs_umap_optim_params optim_params;
    int32_t randomised;          /* accepted for parity with UmapParamsGpu; unused */
    int32_t epoch_strategy;      /* MRS_EPOCH_EDGE_ATOMIC (default) | MRS_EPOCH_NODE_EXACT */
} mrs_umap_params_gpu;

/* ------------------------------------------------------------------ misc */

/* message of the last failure on the calling thread ("" if none) */
const char *mrs_last_error(void);

/* Context: device ordinal, HBM budget in bytes (0 = no limit).  Graphs and
 * optimiser states keep using their context's stream: destroy them before the
 * context. */
int mrs_ctx_create(int32_t device, size_t hbm_budget, mrs_ctx **out);
int mrs_ctx_destroy(mrs_ctx *ctx);
/* block until all work queued on the context's stream has finished */
int mrs_ctx_synchronize(mrs_ctx *ctx);
/* the context's HIP stream (hipStream_t) so callers can order their own work */
void *mrs_ctx_stream(mrs_ctx *ctx);
/* Device memory this context keeps for reuse: every buffer a call allocates
 * comes from the context's cache and returns to it, so repeated fits of one
 * shape allocate nothing after the first.  mrs_ctx_cached_bytes reports the
 * cache's size; mrs_ctx_release_cache waits for the device and returns the
 * cache to it (buffers held by live graphs and states are not affected).  No
 * counterpart in the reference. */
size_t mrs_ctx_cached_bytes(mrs_ctx *ctx);
int mrs_ctx_release_cache(mrs_ctx *ctx);
/* Exhaustive kNN on this context skips database stages that provably hold no
 * candidate (on = 1, the default), or sweeps every stage: in cell order, each
 * query block's own cell first (on = 0), or in row order (on = -1; below
 * 65 536 rows every sweep runs in row order).  All give the same results;
 * the switch exists for comparisons.  No counterpart in the reference. */
int mrs_ctx_set_knn_pruning(mrs_ctx *ctx, int32_t on);
/* The pruned sweep cuts a query block's second-pass stage list longer than
 * `stages` into work items whose partial lists a merge kernel folds back in;
 * 0 (the default) picks max(1024, total / 2048).  The result does not depend
 * on it: smaller values make small problems take the split path, as large
 * ones do by default.  No counterpart in the reference. */
int mrs_ctx_set_knn_work_chunk(mrs_ctx *ctx, int32_t stages);
/* Measurement hook: the rate (gathers per second, written to *per_s) of
 * uniform random 8-byte gathers from a table of `rows` float2 rows, the
 * request pattern of the epoch kernel's negative samples and scattered
 * endpoints.  Times `reps` launches of `gathers` gathers each (8 in flight per
 * thread) on the context's stream after two warm-up launches.  No counterpart
 * in the reference. */
int mrs_probe_random_gather(mrs_ctx *ctx, int64_t rows, int64_t gathers, int32_t reps,
                            double *per_s);
/* The same for FP32 atomic adds without return (the epoch kernel's gradient
 * scatter) to the x of uniform random rows of the table: ops per second.
 * No counterpart in the reference. */
int mrs_probe_random_atomic(mrs_ctx *ctx, int64_t rows, int64_t ops, int32_t reps,
                            double *per_s);

/* parse_umap_optimiser_gpu, src/training/umap_optimiser_gpu.rs:44-52.
 * 0 adam_gpu, 1 adam_parallel, 2 adam, 3 sgd; -1 if unknown. */
int32_t mrs_parse_umap_optimiser_gpu(const char *s);

/* parse_ann_search_gpu, src/data/nearest_neighbours_gpu.rs:203-210.
 * 0 ivf_gpu, 1 nndescent_gpu, 2 exhaustive_gpu; -1 if unknown. */
int32_t mrs_parse_ann_search_gpu(const char *s);

/* Host evaluation of the negative-sample hash the epoch kernels use (same
 * code path, including the invariant-divisor remainder): gpu_hash_neg /
 * cpu_hash_neg, src/training/umap_optimiser_gpu.rs:611-624, :1379-1391.  n >= 1. */
uint32_t mrs_hash_neg(uint32_t seed, uint32_t node, uint32_t epoch, uint32_t edge_local,
                      uint32_t neg, uint32_t n);

/* ------------------------------------------------------------- fits */

/* UmapParamsGpu::default (src/lib.rs:3244-3260) with ann_type "exhaustive_gpu". */
void mrs_umap_params_gpu_default(mrs_umap_params_gpu *p);
/* UmapParamsGpu::new_default_2d (src/lib.rs:3338-3346): the defaults with
 * optim_params = from_min_dist_spread(min_dist, spread); NaN = None (0.5, 1.0).
 * A min_dist/spread pair from_min_dist_spread refuses leaves default_2d's a, b. */
void mrs_umap_params_gpu_default_2d(float min_dist, float spread, mrs_umap_params_gpu *p);

/* UmapOptimParams::from_min_dist_spread (src/training/umap_optimisers.rs:89-122)
 * with fit_params (:139-209): a, b fitted to min_dist and spread by the
 * reference's 300-step gradient descent in f32.  The optional arguments take
 * the reference's defaults when NaN (floats) or negative (ints): lr 1, gamma 1,
 * n_epochs 500, neg_sample_rate 5, beta1 0.5, beta2 0.9, eps 1e-7.  Host only;
 * needs no device.  spread <= 0 or a NaN min_dist is MRS_INVALID_ARGUMENT. */
int mrs_umap_optim_params_from_min_dist_spread(float min_dist, float spread, float lr,
                                               float gamma, int32_t n_epochs,
                                               int32_t neg_sample_rate, float beta1, float beta2,
                                               float eps, mrs_umap_optim_params *out);

/* pca_layout (src/data/init.rs:1065-1140): out [n][n_comp] (host or device)
 * = the first n_comp principal-component scores of data [n][d] (host or
 * device), each column centred and scaled to standard deviation `range` (NaN:
 * PCA_RANGE = 1).  The components are the exact leading eigenvectors of the
 * centred covariance (FP64), for randomised = 0 or 1 alike (the reference's
 * randomised SVD approximates them; seed is accepted for that signature).
 * Each component's sign makes its largest-magnitude loading positive.
 * n_comp 1..min(16, d, n); d <= 2048.  The fit takes this path for
 * initialisation "pca" and, as the reference does, for unknown strings. */
int mrs_pca_layout(mrs_ctx *ctx, const float *data, int64_t n, int32_t d, int32_t n_comp,
                   int32_t randomised, float range, uint64_t seed, float *out);
/* The n_comp leading eigenvectors of the symmetric matrix a [d][d] (FP64) as
 * rows of vecs [n_comp][d] (largest eigenvalue first, each with its
 * largest-magnitude entry positive) and evals [n_comp]: the host eigensolver
 * of the PCA layout (Householder tridiagonalisation, implicit QL).  No
 * counterpart in the reference (faer's SVD); host only, needs no device. */
int mrs_sym_eigen_leading(const double *a, int32_t d, int32_t n_comp, double *vecs,
                          double *evals);

/* umap_gpu (src/lib.rs:3465-3487 -> umap_gpu_inner :3510-3708): data [n*d] f32
 * row-major; optional precomputed kNN (knn_idx/knn_dist [n*knn_k], self
 * excluded, squared distances; pass NULL to search); out [n_dim*n] is the
 * transposed embedding (out[dim*n + i]), as the reference returns [n_dim][n]. */
int mrs_umap_gpu(mrs_ctx *ctx, const float *data, int64_t n, int32_t d, const int32_t *knn_idx,
                 const float *knn_dist, int32_t knn_k, const mrs_umap_params_gpu *params,
                 uint64_t seed, int32_t verbose, float *out);
/* densmap_gpu (src/lib.rs:3829-3851): the same with the density term. */
int mrs_densmap_gpu(mrs_ctx *ctx, const float *data, int64_t n, int32_t d,
                    const int32_t *knn_idx, const float *knn_dist, int32_t knn_k,
                    const mrs_umap_params_gpu *params, const mrs_dens_params *dens,
                    uint64_t seed, int32_t verbose, float *out);
/* Measurement hook (no counterpart in the reference): the calling thread's
 * later fits that search their own kNN copy the lists into idx / dist, device
 * buffers of at least n*k entries (capacity; smaller fits only), with one
 * device-to-device copy on the context's stream.  NULL, NULL switches it off. */
int mrs_fit_knn_copy_to(int32_t *idx, float *dist, int64_t capacity);
/* Stage times of the calling thread's last fit (no counterpart in the reference;
 * its timings are verbose prints): out[0..7] = kNN ms, graph ms, densMAP state
 * ms, init ms, epochs ms, total ms, undirected edges, kNN shortlist-kernel ms. */
int mrs_last_fit_timings(double *out);

/* ------------------------------------------------------------- kNN */

/* Exhaustive kNN of run_ann_search_gpu with "exhaustive_gpu"
 * (src/data/nearest_neighbours_gpu.rs:257-365; the k+1 self query of
 * ann-search-rs 0.5.2 followed by the self filter and take(k) of :350-362).
 * x [n*d] f32 row-major; idx [n*k] int32, dist [n*k] f32 squared Euclidean,
 * ascending, self excluded, ties broken by the smaller index.  precision:
 * MRS_KNN_FP32 or MRS_KNN_BF16_RERANK (MFMA candidates, exact FP32 re-rank of a
 * shortlist of max(shortlist, k) rounded up to 64 or 128 candidates). */
int mrs_knn_exhaustive(mrs_ctx *ctx, const float *x, int64_t n, int32_t d, int32_t k,
                       int32_t precision, int32_t shortlist, int32_t *idx, float *dist);
/* The same search without any allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  x, idx and dist must be
 * device memory; n <= 16384, d <= 128, max(k, shortlist) <= 64 (larger
 * problems are MRS_INVALID_ARGUMENT).  Results equal mrs_knn_exhaustive's bit
 * for bit.  The work is queued on the context's stream and the call returns
 * without waiting for it (mrs_ctx_synchronize, or order later work on
 * mrs_ctx_stream).  Its working memory is one 72 MiB device array per GPU,
 * kept until the process ends and shared by every call on that GPU: calls
 * that overlap on one GPU (other contexts' streams) race. */
int mrs_knn_exhaustive_noalloc(mrs_ctx *ctx, const float *x, int64_t n, int32_t d, int32_t k,
                               int32_t precision, int32_t shortlist, int32_t *idx, float *dist);
/* Same, returning in ms_out[0..9]: the stage times in ms of pack (with the
 * cell layout when pruning), shortlist (every pass), merge+re-rank; database
 * splits used; shortlist size; the fraction of (query block, database stage)
 * pairs swept (1 without pruning); cells (0 without pruning); the longest and
 * the mean second-pass stage list of a query block; the second-pass work items
 * split off a longer list.  No counterpart in the reference; used by the
 * benchmark and the tests. */
int mrs_knn_exhaustive_timed(mrs_ctx *ctx, const float *x, int64_t n, int32_t d, int32_t k,
                             int32_t precision, int32_t shortlist, int32_t *idx, float *dist,
                             double *ms_out);

/* ------------------------------------------------------------ graph */

/* The graph half of construct_umap_graph_gpu (src/lib.rs:3376-3441) on the
 * device: smooth_knn_dist (src/data/graph.rs:499-565, bandwidth search in FP64),
 * knn_to_coo (:585-662), symmetrise_graph (:684-694) and filter_weak_edges
 * (:776-853), plus the epoch layout of UmapCsrGraph::from_graph
 * (src/training/umap_optimiser_gpu.rs:132-209).  knn_idx/knn_dist are [n*k]
 * (self excluded, ascending squared distances), host or device memory. */
int mrs_fuzzy_graph(mrs_ctx *ctx, const int32_t *knn_idx, const float *knn_dist, int64_t n,
                    int32_t k, const mrs_umap_graph_params *params, int32_t n_epochs,
                    mrs_graph **out);
/* A graph from a caller's CoordinateList (src/data/structures.rs:22-31): the
 * symmetric fuzzy graph symmetrise_graph / filter_weak_edges produce
 * (src/data/graph.rs:684-694, :776-853), entries in any order, host or device
 * pointers.  It is filtered again with n_epochs (which keeps an already
 * filtered graph as it is) and laid out as mrs_fuzzy_graph does, so
 * mrs_graph_export_coo returns it in (row, col) order.  Entries out of range,
 * on the diagonal, repeated, or without their mirror (c, r) are rejected with
 * MRS_INVALID_ARGUMENT.  The graph has no kNN behind it: mrs_graph_sigma_rho
 * and mrs_dens_state_from_graph refuse it. */
int mrs_graph_from_coo(mrs_ctx *ctx, int64_t n, const int32_t *rows, const int32_t *cols,
                       const float *vals, int64_t nnz, int32_t n_epochs, mrs_graph **out);
int mrs_graph_destroy(mrs_graph *g);
/* n, number of COO entries after filtering (both directions), undirected edges */
int mrs_graph_info(const mrs_graph *g, int64_t *n, int64_t *nnz, int64_t *n_edges);
/* the filtered CoordinateList (src/data/structures.rs:22-31): row-major, columns
 * ascending; buffers of length nnz (host or device) */
int mrs_graph_export_coo(const mrs_graph *g, int32_t *rows, int32_t *cols, float *vals);
/* sigma and rho of smooth_knn_dist, [n] each */
int mrs_graph_sigma_rho(const mrs_graph *g, float *sigma, float *rho);
/* DensState::new (src/utils/density.rs:181-201) from the device graph and the
 * kNN it was built from: r [n], mu_sum [n], mu_tot. */
int mrs_dens_state_from_graph(mrs_ctx *ctx, const mrs_graph *g, const int32_t *knn_idx,
                              const float *knn_dist, int32_t distances_squared, float *r_out,
                              float *mu_sum_out, float *mu_tot_out);

/* spectral_layout (src/data/init.rs:986-1004) of a device graph: out [n*n_dim]
 * row-major (host or device).  range NaN = SPECTRAL_RANGE (10).  data [n*d]
 * (optional, NULL allowed) places component centroids on the disconnected path
 * as the reference's feature matrix does (:539-595). */
int mrs_spectral_layout(mrs_ctx *ctx, const mrs_graph *g, int32_t n_dim, uint64_t seed,
                        float range, const float *data, int32_t d, float *out);

/* ------------------------------------------------- optimiser (adam_gpu) */

/* UmapGpuState::upload (src/training/umap_optimiser_gpu.rs:374-452) preceded by
 * UmapCsrGraph::from_graph (:132-209).  The graph is an adjacency list in CSR
 * form: graph[i] = (adj_cols, adj_vals)[adj_off[i] .. adj_off[i+1]).
 * embd is [n * n_dim] row-major.  dens may be NULL (plain UMAP). */
int mrs_umap_state_create(mrs_ctx *ctx, const float *embd, int64_t n, int32_t n_dim,
                          const int64_t *adj_off, const int32_t *adj_cols,
                          const float *adj_vals, const mrs_umap_optim_params *params,
                          const mrs_dens_state *dens, mrs_umap_state **out);
/* The same state built from a device graph (no host adjacency list): the
 * device-side equivalent of coo_to_adjacency_list + UmapCsrGraph::from_graph +
 * upload (src/lib.rs:3612, umap_optimiser_gpu.rs:132-209, :374-452). */
int mrs_umap_state_from_graph(mrs_ctx *ctx, const mrs_graph *g, const float *embd,
                              int32_t n_dim, const mrs_umap_optim_params *params,
                              const mrs_dens_state *dens, mrs_umap_state **out);
int mrs_umap_state_destroy(mrs_umap_state *st);

/* sizes of a state: n, n_dim, n_edges */
int mrs_umap_state_info(const mrs_umap_state *st, int64_t *n, int32_t *n_dim,
                        int64_t *n_edges);

/* Read/write a state buffer to/from caller memory.  which:
 *  0 embd [n*dim] f32     1 m [n*dim]          2 v [n*dim]
 *  3 node_grad [n*dim]    4 has_update [n] u32
 *  5 node_edge_offsets [n+1] u32  6 csr_edge_idx [2E] u32  7 csr_other_node [2E] u32
 *  8 epochs_per_sample [E] f32    9 epoch_of_next_sample [E] f32   10 mu [E] f32
 * 11 dens_re [n]  12 dens_phi_sum [n]  13 dens_weight [n]  14 dens_r [n]
 * 15 edge endpoints [2E] u32 (i, j interleaved, i < j, edge order; read-only)
 * 16 edge CSR slots [2E] u32 (the edge's position in row i and in row j,
 *    interleaved: the edge_local keys of its negative samples; read-only) */
int mrs_umap_state_read(mrs_umap_state *st, int32_t which, void *out);
int mrs_umap_state_write(mrs_umap_state *st, int32_t which, const void *in);

/* launch_grad_accum, umap_optimiser_gpu.rs:1016-1068 (node kernel :682-860).
 * dens_coeff == NULL selects the plain-UMAP arm. */
int mrs_launch_grad_accum(mrs_umap_state *st, int32_t epoch, uint32_t seed,
                          const float *dens_coeff);
/* launch_dens_update, umap_optimiser_gpu.rs:1089-1143: radii, FP64 statistics
 * and weights, all on the device; returns dens_coeff. */
int mrs_launch_dens_update(mrs_umap_state *st, float *dens_coeff_out);
/* launch_adam_update, umap_optimiser_gpu.rs:1155-1196 */
int mrs_launch_adam_update(mrs_umap_state *st, int32_t epoch);
/* launch_edge_schedule_update, umap_optimiser_gpu.rs:1205-1227 */
int mrs_launch_edge_schedule_update(mrs_umap_state *st, int32_t epoch);
/* launch_grad_norm, umap_optimiser_gpu.rs:1245-1272 */
int mrs_launch_grad_norm(mrs_umap_state *st, float *norm_out);

/* Edge-parallel counterpart of launch_grad_accum + launch_edge_schedule_update:
 * one thread per undirected edge, FP32 atomic scatter-add of the gradient into
 * node_grad and has_update, schedule cursors advanced in place.  node_grad and
 * has_update are cleared first if a node-parallel launch left them dirty.
 * dens_coeff == NULL selects the plain-UMAP arm; otherwise the weights and radii
 * of the last mrs_launch_dens_update are used with this coefficient. */
int mrs_launch_epoch_edges(mrs_umap_state *st, int32_t epoch, uint32_t seed,
                           const float *dens_coeff);
/* Adam step (launch_adam_update semantics) for flagged nodes that also clears
 * node_grad and has_update, closing an edge-parallel epoch. */
int mrs_launch_adam_clear(mrs_umap_state *st, int32_t epoch);

/* The epoch loop of optimise_embedding_adam_gpu (umap_optimiser_gpu.rs:1336-1360)
 * for epochs [first, first+count), with the chosen strategy
 * (MRS_EPOCH_EDGE_ATOMIC or MRS_EPOCH_NODE_EXACT).  Asynchronous on the
 * context stream; a density epoch synchronises nothing (all statistics stay on
 * the device). */
int mrs_run_epochs(mrs_umap_state *st, int32_t first, int32_t count, uint64_t seed,
                   int32_t strategy);

/* Per-kernel timing of mrs_run_epochs with HIP events on the context stream
 * (no counterpart in the reference; used by the benchmark).  enable=1 starts
 * recording, enable=0 stops; mrs_umap_state_kernel_times waits for the
 * recorded work and returns accumulated milliseconds and launch counts per
 * kind: 0 edge epoch, 1 adam+clear, 2 node gradient, 3 adam, 4 schedule,
 * 5 densMAP update, 6 firing bitmask of a window of edge-parallel epochs
 * (arrays of 8). */
int mrs_umap_state_set_profiling(mrs_umap_state *st, int32_t enable);
int mrs_umap_state_kernel_times(mrs_umap_state *st, double *ms_out, int64_t *count_out);

/* Reset a state to the start of a fit (UmapGpuState::upload's initial values,
 * umap_optimiser_gpu.rs:411-441): embd from the caller ([n*dim], host or device),
 * m = v = node_grad = 0, has_update = 0, epoch_of_next_sample = epochs_per_sample. */
int mrs_umap_state_reset(mrs_umap_state *st, const float *embd);

/* Number of edges that tick in each epoch of the schedule
 * (epoch_of_next_sample <= epoch, umap_optimiser_gpu.rs:940-955), replayed on the
 * host in f32 from epochs_per_sample [n_edges]; counts_out is [n_epochs].  The
 * sum is the fit's total edge-updates. */
int mrs_schedule_active_counts(const float *epochs_per_sample, int64_t n_edges,
                               int32_t n_epochs, int64_t *counts_out);

/* optimise_embedding_adam_gpu, src/training/umap_optimiser_gpu.rs:1302-1369.
 * embd [n*n_dim] is updated in place. */
int mrs_optimise_embedding_adam_gpu(mrs_ctx *ctx, float *embd, int64_t n, int32_t n_dim,
                                    const int64_t *adj_off, const int32_t *adj_cols,
                                    const float *adj_vals,
                                    const mrs_umap_optim_params *params,
                                    const mrs_dens_state *dens, uint64_t seed,
                                    int32_t verbose, int32_t strategy);

/* Host solver of the Lanczos tridiagonal problems of the spectral layout
 * (compute_smallest_eigenpairs_lanczos's eigen step, src/utils/math.rs:560-590):
 * symmetric tridiagonal diag [m], offdiag [m-1] -> eigenvalues ascending [m]
 * and eigenvectors as columns of evecs [m][m] (row-major).  Host memory only;
 * needs no device. */
int mrs_tridiag_eigen(const double *diag, const double *offdiag, int32_t m, double *evals,
                      double *evecs);

/* ---------------------------------------------------------------- serving
 * A fit server: continuous batching of many uneven umap_gpu fits on one GPU.
 * The reference has no serving layer; each request has exactly the semantics
 * of mrs_umap_gpu (src/lib.rs:3465-3487) with the same parameters and seed.
 * Requests are admitted in arrival order while their estimated device memory
 * (mrs_fit_bytes_estimate) fits the HBM budget and fewer than max_active fits
 * are resident; prep_workers threads prepare admitted fits concurrently (kNN,
 * graph, spectral layout, each worker on its own stream); one batcher advances
 * every resident fit by one epoch per tick (one segmented launch pair for all
 * edge-parallel fits, per-fit launches for node-exact ones).  Node-exact
 * requests return exactly what mrs_umap_gpu returns; edge-parallel ones differ
 * from it only by the FP32 atomic summation order. */
typedef struct mrs_server mrs_server;

/* Upper estimate of one fit's peak device bytes (preparation and resident state). */
size_t mrs_fit_bytes_estimate(int64_t n, int32_t d, int32_t k, int32_t n_dim);

/* hbm_budget 0 = half the device's free memory at creation.  prep_workers
 * 1..8 (larger counts are refused: a run with 12 workers ended in a device
 * memory fault whose cause has not been found); max_active 1..4096. */
int mrs_server_create(int32_t device, size_t hbm_budget, int32_t prep_workers, int32_t max_active,
                      mrs_server **out);

/* Queue one fit of data [n][d] (host or device; it must stay valid until the
 * request completes).  params as for mrs_umap_gpu (densMAP is not served).
 * A request larger than the whole budget is InvalidArgument. */
int mrs_server_submit(mrs_server *sv, const float *data, int64_t n, int32_t d,
                      const mrs_umap_params_gpu *params, uint64_t seed, int64_t *ticket);

/* state: 0 queued, 1 preparing, 2 in epochs, 3 done, 4 failed. */
int mrs_server_poll(mrs_server *sv, int64_t ticket, int32_t *state);

/* Block until the request completes; copy its [n_dim][n] embedding to out
 * (host or device; may be NULL) and return the request's status.
 * times_ms (may be NULL) = {queued, preparing, submit-to-result}.  The ticket
 * is consumed. */
int mrs_server_wait(mrs_server *sv, int64_t ticket, float *out, double *times_ms);

/* out[7] = {submitted, completed, failed, ticks, peak resident fits,
 *           peak admitted bytes, budget bytes}. */
int mrs_server_stats(mrs_server *sv, int64_t *out);

/* Finishes every queued request, then stops the threads and frees everything. */
int mrs_server_destroy(mrs_server *sv);

/* ------------------------------------------------------------- autotuning
 * Launch shapes chosen by measurement.  No counterpart in the reference,
 * whose workgroup sizes are compile-time constants
 * (src/training/umap_optimiser_gpu.rs:72-82).  Two families have several
 * valid shapes with the same results:
 *   MRS_TUNE_EPOCH_SPAN  edges per wavefront of the edge-parallel epoch
 *                        (64, 128, 256); only the order of the FP32 atomic
 *                        additions differs between them;
 *   MRS_TUNE_KNN_SPLITS  database splits of the unpruned kNN sweep (1, 2, 4,
 *                        8); the lists are bit-identical.
 * With MRS_TUNE_ON_FIRST_USE (the default) the first call that meets an
 * undecided shape class (dimension, density term, and edge and node counts
 * as powers of two; precision, padded width, shortlist and rows as a power
 * of two)
 * times every candidate on its own data with HIP events and records the
 * fastest: 12 epochs in the middle of a run's stretch with one density flag
 * (at least 12 epochs long) alternate the spans; the first search runs the
 * shortlist once per split count.  Later
 * calls of the class use the record.  MRS_TUNE_OFF keeps the built-in
 * choices (spans of 128 from 2^20 edges, else 64; splits filling the chip
 * with 1024 workgroups).  Process-wide; thread-safe. */
#define MRS_TUNE_OFF 0
#define MRS_TUNE_ON_FIRST_USE 1
#define MRS_TUNE_EPOCH_SPAN 1
#define MRS_TUNE_KNN_SPLITS 2
int mrs_autotune_set_mode(int32_t mode);
/* Use `variant` for every call of the family (overrides records and mode);
 * -1 releases it.  For tests and experiments. */
int mrs_autotune_force(int32_t family, int32_t variant);
/* Forget every recorded decision. */
int mrs_autotune_clear(void);
/* The records as text, one per line: "family shape variant cand:ms ...".
 * Writes at most cap bytes (NUL-terminated) and returns the size needed. */
int64_t mrs_autotune_dump(char *buf, int64_t cap);
/* Save / load the records (text; the first line names the device they were
 * measured on, and a table from another device is refused). */
int mrs_autotune_save(const char *path);
int mrs_autotune_load(const char *path);

#ifdef __cplusplus
}
#endif
#endif
