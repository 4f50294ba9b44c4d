/*
 * oracle.h -- CPU restatement of the manifolds-rs UMAP/densMAP fitting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path and the timed CPU baseline ("port").  Only tests/, the smoke() check
 * and the cpu_baseline leg of bench.py may load it; the product library never
 * links or calls it.
 *
 * Every function restates one reference function; the citation
 * (path:line into GregorLueg/manifolds-rs v0.3.15) sits above each declaration.
 * Arithmetic is done in the reference's T = f32 unless noted, with FMA
 * contraction disabled (Rust never fuses), so results are bit-comparable.
 *
 * Pinning status per function is recorded in oracle.c's header comment.
 */
#ifndef MRS_ORACLE_H
#define MRS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- optimiser constants: src/training/umap_optimiser_gpu.rs:58-70 --- */
#define ORC_GRAD_CLIP_VAL 4.0f
#define ORC_GRAD_REP_EPS 0.001f
#define ORC_GRAD_DIST_SQ_THRESHOLD 1e-8f
/* src/utils/density.rs:38 and :61 */
#define ORC_DENS_EPS 1e-8
#define ORC_DENS_REDUCE_CHUNK 8192

typedef struct {
    float a, b, lr, gamma;
    int32_t n_epochs, neg_sample_rate;
    float min_dist, beta1, beta2, eps;
} orc_optim_params; /* UmapOptimParams<f32>, src/training/umap_optimisers.rs:22-44 */

/* splitmix-style negative hash: umap_optimiser_gpu.rs:611-624 (= :1379-1391) */
uint32_t orc_hash_neg(uint32_t seed, uint32_t node, uint32_t epoch, uint32_t edge_local,
                      uint32_t neg, uint32_t n);

/* UmapCsrGraph::from_graph, umap_optimiser_gpu.rs:132-209.
 * Input adjacency graph[i] = adj_cols/adj_vals[adj_off[i]..adj_off[i+1]].
 * Pass NULL outputs to only count edges (returned in *n_edges).
 * Returns 0 ok, 1 NoData, 2 NoGraphEdges. */
int orc_csr_from_graph(int64_t n, const int64_t *adj_off, const int32_t *adj_cols,
                       const float *adj_vals, int64_t *n_edges, uint32_t *node_edge_offsets,
                       uint32_t *csr_edge_idx, uint32_t *csr_other_node,
                       float *epochs_per_sample, float *mu, int32_t *edge_i, int32_t *edge_j);

/* cpu_grad_accum, umap_optimiser_gpu.rs:1461-1579 (GPU kernel semantics :682-860).
 * dens_* may be NULL (plain UMAP).  has_update may be NULL.  abs_out (NULL
 * allowed) receives per coordinate the sum of the absolute values of the terms
 * the gradient adds up, the scale of its rounding error. */
void orc_grad_accum(const float *embd, int64_t n, int32_t n_dim, const uint32_t *off,
                    const uint32_t *csr_edge_idx, const uint32_t *csr_other,
                    const float *next_sample, int32_t epoch, uint32_t seed, int32_t neg_rate,
                    const orc_optim_params *p, const float *mu, const float *dens_re,
                    const float *dens_phi_sum, const float *dens_weight, float dens_coeff,
                    float *grad_out, uint32_t *has_update, float *abs_out);

/* the same terms summed in double and rounded once: the exact sum of the
 * reference's f32 terms, against which any summation order is measured */
void orc_grad_accum_exact(const float *embd, int64_t n, int32_t n_dim, const uint32_t *off,
                          const uint32_t *csr_edge_idx, const uint32_t *csr_other,
                          const float *next_sample, int32_t epoch, uint32_t seed,
                          int32_t neg_rate, const orc_optim_params *p, const float *mu,
                          const float *dens_re, const float *dens_phi_sum,
                          const float *dens_weight, float dens_coeff, float *grad_out,
                          uint32_t *has_update, float *abs_out);

/* per-epoch Adam scalars, umap_optimiser_gpu.rs:1155-1196 / :1595-1604 */
void orc_adam_scalars(const orc_optim_params *p, int32_t epoch, int32_t n_epochs,
                      float *fused_lr, float *epsc);

/* cpu_adam_update, umap_optimiser_gpu.rs:1584-1627 */
void orc_adam_update(const float *grad, const uint32_t *has_update, float *m, float *v,
                     float *embd, int64_t n, int32_t n_dim, const orc_optim_params *p,
                     int32_t epoch, int32_t n_epochs);

/* umap_edge_schedule_update, umap_optimiser_gpu.rs:940-955 */
void orc_schedule_update(const float *per, float *next, int64_t n_edges, int32_t epoch);

/* cpu_dens_radii, umap_optimiser_gpu.rs:1409-1453 */
void orc_dens_radii(const float *embd, int64_t n, int32_t n_dim, const uint32_t *off,
                    const uint32_t *csr_other, float a, float b, float *re, float *phi_sum);

/* correlation_stats, src/utils/density.rs:464-508 (f64, fixed 8192 chunks).
 * out = {mean, std, cov} as the reference's f64 values before the cast to T. */
void orc_correlation_stats(const float *re, const float *r, int64_t n, float var_shift,
                           double out[3]);

/* launch_dens_update host side, umap_optimiser_gpu.rs:1122-1142: weights and coeff */
float orc_dens_weights(const float *r, const float *re, int64_t n, float var_shift,
                       float lambda, float mu_tot, float *weight);

/* DensState::is_active, src/utils/density.rs:244-250 (T = f32) */
int orc_dens_is_active(float lambda, float frac, int32_t epoch, int32_t n_epochs);

/* optimise_embedding_adam_gpu driver replay on the CPU, umap_optimiser_gpu.rs:1302-1369.
 * dens_r / mu_tot / dens params may be NULL/0 for plain UMAP.  seed is truncated to u32. */
int orc_optimise_adam_gpu(float *embd, int64_t n, int32_t n_dim, const int64_t *adj_off,
                          const int32_t *adj_cols, const float *adj_vals,
                          const orc_optim_params *p, const float *dens_r, float dens_lambda,
                          float dens_frac, float dens_var_shift, float dens_mu_tot,
                          uint64_t seed);

/* smooth_knn_dist + local_connectivity_distance, src/data/graph.rs:107-125, 499-565 */
void orc_smooth_knn_dist(const float *dist, int64_t n, int32_t k, float local_connectivity,
                         float bandwidth, int32_t n_iter, float *sigma, float *rho);

/* knn_to_coo, src/data/graph.rs:585-662.  Writes up to n*k entries, returns count. */
int64_t orc_knn_to_coo(const int32_t *idx, const float *dist, int64_t n, int32_t k,
                       const float *sigma, const float *rho, int32_t *rows, int32_t *cols,
                       float *vals);

/* symmetrise_graph -> symmetrise_csr, src/data/graph.rs:185-468, 684-694.
 * Output buffers must hold 2*nnz entries.  Returns output count. */
int64_t orc_symmetrise(int64_t n, const int32_t *rows, const int32_t *cols, const float *vals,
                       int64_t nnz, float mix_weight, int32_t *out_rows, int32_t *out_cols,
                       float *out_vals);

/* filter_weak_edges, src/data/graph.rs:776-853.  In-place compaction; returns count. */
int64_t orc_filter_weak_edges(int32_t *rows, int32_t *cols, float *vals, int64_t nnz,
                              int32_t n_epochs);

/* Exhaustive kNN (self excluded, ascending, ties by index) -- the textbook
 * brute force the un-vendored ann-search-rs exhaustive index is specified by
 * (call sites src/data/nearest_neighbours_gpu.rs:282-290, self filter :350-362).
 * Squared Euclidean, accumulated with fmaf over dims in order.  Rows
 * [q0, q1) are computed (OpenMP).  f64 variant computes in double. */
void orc_knn_exact_f32(const float *x, int64_t n, int32_t d, int32_t k, int64_t q0, int64_t q1,
                       int32_t *idx, float *dist);
void orc_knn_exact_f64(const float *x, int64_t n, int32_t d, int32_t k, int64_t q0, int64_t q1,
                       int32_t *idx, double *dist);

/* the f64 brute force for an arbitrary list of query rows (OpenMP over rows) */
void orc_knn_exact_f64_rows(const float *x, int64_t n, int32_t d, int32_t k,
                            const int64_t *rows, int64_t n_rows, int32_t *idx, double *dist);

/* edge_distances_from_knn + original_log_radii + standardise + mu_tot,
 * src/utils/density.rs:181-201, 309-438.  rows must be grouped as the COO is.
 * Returns 0 ok, 1 NoData, 3 DegenerateLocalRadii. */
int orc_dens_state(int64_t n, const int32_t *rows, const int32_t *cols, const float *vals,
                   int64_t nnz, const int32_t *knn_idx, const float *knn_dist, int32_t k,
                   int32_t distances_squared, float *r_out, float *mu_sum_out,
                   float *mu_tot_out);

/* UmapOptimParams::fit_params, src/training/umap_optimisers.rs:139-209 (T = f32) */
void orc_fit_params(float min_dist, float spread, int32_t n_iter, float *a, float *b);
/* one iteration's normalised (grad_a, grad_b) of fit_params at (a, b), :170-193 */
void orc_fit_params_grad(float min_dist, float spread, float a, float b, float *ga, float *gb);

/* FastPowLut::new(b, 25.0, 65_536).get(x) for each x, src/training/umap_optimisers.rs:325-397 */
void orc_lut_pow(float b, const float *x, int64_t n, float *out);

/* density_edge_coeff, src/training/umap_optimisers.rs:564-592 (phi_sum and mu
 * floored at FLT_EPSILON) */
float orc_density_edge_coeff(float a, float b, float dist_sq, float dist_sq_b, float denom,
                             float phi_sum_i, float phi_sum_o, float re_i, float re_o, float w_i,
                             float w_o, float coeff, float mu_edge);

/* optimise_embedding_adam_parallel, src/training/umap_optimisers.rs:1170-1492.
 * Negatives: the reference draws from per-node SmallRng streams (parity
 * unpinned); here they are injected from orc_hash_neg(seed, node, epoch, slot, s).
 * use_lut selects the reference's 65 536-entry pow LUT (:325-397). */
int orc_optimise_adam_parallel(float *embd, int64_t n, int32_t n_dim, const int64_t *adj_off,
                               const int32_t *adj_cols, const float *adj_vals,
                               const orc_optim_params *p, uint64_t seed, int32_t use_lut,
                               int32_t n_threads);

/* the same with the densMAP term (dens_r NULL = plain UMAP) */
int orc_optimise_adam_parallel_dens(float *embd, int64_t n, int32_t n_dim,
                                    const int64_t *adj_off, const int32_t *adj_cols,
                                    const float *adj_vals, const orc_optim_params *p,
                                    uint64_t seed, int32_t use_lut, int32_t n_threads,
                                    const float *dens_r, float dens_lambda, float dens_frac,
                                    float dens_var_shift, float dens_mu_tot);

/* optimise_embedding_sgd, src/training/umap_optimisers.rs:653-865 (plain UMAP,
 * negatives injected from the hash as above). */
int orc_optimise_sgd(float *embd, int64_t n, int32_t n_dim, const int64_t *adj_off,
                     const int32_t *adj_cols, const float *adj_vals, const orc_optim_params *p,
                     uint64_t seed);

/* spectral_layout, src/data/init.rs:986-1004 (components, Lanczos with
 * modified Gram-Schmidt, src/utils/math.rs:401-627, meta layouts for
 * disconnected graphs, centring, scaling, noise).  A symmetric COO of n
 * vertices; data [n][d] or NULL; out [n][n_comp].  StdRng draws replaced by
 * counter-based draws of the same distributions.  0 ok, 1 NoData. */
int orc_spectral_layout(int64_t n, const int32_t *rows, const int32_t *cols, const float *vals,
                        int64_t nnz, int32_t n_comp, uint64_t seed, float range, const float *data,
                        int32_t d, float *out);
/* The same with each eigenvector's free sign (components, meta layout) chosen
 * to agree with align [n][n_comp], another layout of the same graph (NULL:
 * none).  Test use: compares a layout up to those signs. */
int orc_spectral_layout_aligned(int64_t n, const int32_t *rows, const int32_t *cols,
                                const float *vals, int64_t nnz, int32_t n_comp, uint64_t seed,
                                float range, const float *data, int32_t d, const float *align,
                                float *out);

/* pca_layout, src/data/init.rs:1065-1140: x [n][d], out [n][n_comp]; the
 * leading eigenvectors of the centred covariance by cyclic Jacobi (FP64),
 * each with its largest-magnitude entry positive; every score column centred
 * and scaled to standard deviation `range`.  0 ok, 1 NoData, 5 bad n_comp. */
int orc_pca_layout(const float *x, int64_t n, int32_t d, int32_t n_comp, float range, float *out);

#ifdef __cplusplus
}
#endif
#endif
