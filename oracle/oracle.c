/*
 * oracle.c -- CPU restatement of the manifolds-rs UMAP/densMAP fitting path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline
 * ("port").  Loaded by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg; never linked into or called by the product library.
 *
 * Built with -O3 -ffp-contract=off (no fused multiply-add unless written as
 * fmaf), so float arithmetic follows the reference's f32 operation order.
 *
 * Pinning (what fixes each function independently of its own formula):
 *   orc_hash_neg            the finaliser is C. Wellons' lowbias32 (checked with
 *                           its published inverse); the pre-mix multipliers are
 *                           the xxHash32 primes; reduction = plain % n
 *   orc_csr_from_graph      reference golden triangle layout / periods / errors
 *                           (umap_optimiser_gpu.rs:1716-1785)
 *   orc_grad_accum          hand-computed (1.2,0,-1.2,0) case (:2060-2092), empty
 *                           epoch = 0 (:1995-2003), finite differences of the UMAP
 *                           loss for the attractive term, brute-force restatement
 *                           in float64 for the repulsive term
 *   orc_adam_update         hand-computed m=0.1, v=0.001, y=1 (:2211-2242),
 *                           inactive node preserved (:2192-2209), torch.optim-free
 *                           closed form of bias-corrected Adam
 *   orc_schedule_update     golden schedules (:2342-2379)
 *   orc_dens_radii          closed form on a 2-node graph
 *   orc_correlation_stats   reference known values (density.rs:651-676)
 *   orc_dens_state          reference known value (density.rs:592-649)
 *   orc_dens_is_active      reference cases (density.rs:691-718)
 *   orc_smooth_knn_dist     invariant |sum exp - ln k| < tol (or 64 iterations),
 *                           rho = d[0] (graph.rs:1441-1471, tests/umap_tests.rs)
 *   orc_knn_to_coo          closed form, self excluded (graph.rs:1473-1510)
 *   orc_symmetrise          goldens 0.86/0.76/0.92 (graph.rs:1512-1563) and a
 *                           dense-matrix restatement on random small graphs
 *   orc_filter_weak_edges   closed-form threshold on random inputs
 *   orc_knn_exact_*         scipy cdist + lexsort brute force on small inputs
 *   orc_fit_params          bound/shape invariants (umap_optimisers.rs:1559-1593);
 *                           its step (orc_fit_params_grad) equals central
 *                           differences of the stated mean squared curve error;
 *                           n_iter steps compose as projected gradient descent
 *   orc_lut_pow             x^b within the linear-interpolation error bound of
 *                           each cell (h^2/8 max|f''|), exact powf from 25 up
 *   orc_density_edge_coeff  hand values at a = b = d^2 = 1 and its max(eps)
 *                           floors on a zero phi sum and a zero edge weight
 *   orc_optimise_adam_gpu   composition of the pinned pieces above
 *   orc_optimise_adam_parallel, orc_optimise_sgd
 *                           hand-computed float64 goldens on 2- and 3-node graphs
 *                           (schedule, first tick drawing rate - 1 negatives,
 *                           floor of fractional negative counts, negatives on the
 *                           smaller endpoint only, in-place SGD with the
 *                           attractive clip, the densMAP term); the reference's
 *                           own negatives come from rand 0.9.5 SmallRng streams
 *                           that cannot be run here, so these functions draw the
 *                           hash's negatives instead (the GPU path's keys).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ hash */

/* umap_optimiser_gpu.rs:611-624 */
uint32_t orc_hash_neg(uint32_t seed, uint32_t node, uint32_t epoch, uint32_t edge_local,
                      uint32_t neg, uint32_t n) {
    uint32_t h = seed ^ (node * 0x9E3779B1u) ^ (epoch * 0x85EBCA77u) ^
                 (edge_local * 0xC2B2AE3Du) ^ (neg * 0x27D4EB2Fu);
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h % n;
}

/* ------------------------------------------------------------------- CSR */

/* UmapCsrGraph::from_graph, umap_optimiser_gpu.rs:132-209 */
int orc_csr_from_graph(int64_t n, const int64_t *adj_off, const int32_t *adj_cols,
                       const float *adj_vals, int64_t *n_edges_out,
                       uint32_t *node_edge_offsets, uint32_t *csr_edge_idx,
                       uint32_t *csr_other_node, float *epochs_per_sample, float *mu,
                       int32_t *edge_i, int32_t *edge_j) {
    if (n == 0) return 1;
    int64_t ne = 0;
    for (int64_t i = 0; i < n; i++)
        for (int64_t p = adj_off[i]; p < adj_off[i + 1]; p++)
            if (i < adj_cols[p]) ne++;
    *n_edges_out = ne;
    if (ne == 0) return 2;
    if (!node_edge_offsets) return 0;

    uint32_t *degree = (uint32_t *)calloc((size_t)n, sizeof(uint32_t));
    float max_w = 0.0f;
    int64_t e = 0;
    for (int64_t i = 0; i < n; i++)
        for (int64_t p = adj_off[i]; p < adj_off[i + 1]; p++) {
            int64_t j = adj_cols[p];
            if (i < j) {
                degree[i]++;
                degree[j]++;
                float w = adj_vals[p];
                if (w > max_w) max_w = w;
                if (edge_i) edge_i[e] = (int32_t)i;
                if (edge_j) edge_j[e] = (int32_t)j;
                e++;
            }
        }
    node_edge_offsets[0] = 0;
    for (int64_t i = 0; i < n; i++) node_edge_offsets[i + 1] = node_edge_offsets[i] + degree[i];
    uint32_t *cursor = degree; /* reuse */
    for (int64_t i = 0; i < n; i++) cursor[i] = node_edge_offsets[i];
    e = 0;
    for (int64_t i = 0; i < n; i++)
        for (int64_t p = adj_off[i]; p < adj_off[i + 1]; p++) {
            int64_t j = adj_cols[p];
            if (i < j) {
                uint32_t pi = cursor[i]++;
                csr_edge_idx[pi] = (uint32_t)e;
                csr_other_node[pi] = (uint32_t)j;
                uint32_t pj = cursor[j]++;
                csr_edge_idx[pj] = (uint32_t)e;
                csr_other_node[pj] = (uint32_t)i;
                float w = adj_vals[p];
                float norm = w / max_w;
                epochs_per_sample[e] = norm > 0.0f ? 1.0f / norm : 1e8f;
                mu[e] = w;
                e++;
            }
        }
    free(degree);
    return 0;
}

/* ------------------------------------------------------- gradient mirror */

/* cpu_grad_accum, umap_optimiser_gpu.rs:1461-1579.  exact != 0: the same f32
 * terms, summed in double and rounded once (the exact sum of the reference's
 * terms, the yardstick for any summation order) */
static void grad_accum_impl(const float *embd, int64_t n, int32_t n_dim, const uint32_t *off,
                            const uint32_t *csr_edge_idx, const uint32_t *csr_other,
                            const float *next_sample, int32_t epoch, uint32_t seed,
                            int32_t neg_rate, const orc_optim_params *p, const float *mu,
                            const float *dens_re, const float *dens_phi_sum,
                            const float *dens_weight, float dens_coeff, float *grad,
                            uint32_t *has_update, float *abs_out, int exact) {
    const float two = 2.0f, one = 1.0f;
    const float clip = ORC_GRAD_CLIP_VAL, rep_eps = ORC_GRAD_REP_EPS;
    const float two_a_b = two * p->a * p->b;
    const float two_gamma_b = two * p->gamma * p->b;
    const float epoch_t = (float)epoch;
    const int dens = dens_re != NULL;
    memset(grad, 0, sizeof(float) * (size_t)n * (size_t)n_dim);
    if (abs_out) memset(abs_out, 0, sizeof(float) * (size_t)n * (size_t)n_dim);
    /* nodes are independent (each writes only its own gradient): parallel over nodes */
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t node = 0; node < n; node++) {
        const float *ys = embd + node * n_dim;
        float *g = grad + node * n_dim;
        double gx[16] = {0}; /* the exact mode's sums (n_dim <= 16) */
        /* |term| sums per coordinate, the scale of each gradient's rounding */
        float *ab = abs_out ? abs_out + node * n_dim : NULL;
        uint32_t active = 0, edge_local = 0;
#define ADD_TERM(d, t)                   \
    do {                                 \
        float term_ = (t);               \
        if (exact) gx[d] += (double)term_; \
        else g[d] += term_;              \
        if (ab) ab[d] += fabsf(term_);   \
    } while (0)
        for (uint32_t pos = off[node]; pos < off[node + 1]; pos++, edge_local++) {
            uint32_t e = csr_edge_idx[pos];
            if (next_sample[e] > epoch_t) continue;
            active = 1;
            uint32_t other = csr_other[pos];
            const float *yo = embd + (int64_t)other * n_dim;
            float dist_sq = 0.0f;
            for (int d = 0; d < n_dim; d++) {
                float diff = ys[d] - yo[d];
                dist_sq += diff * diff;
            }
            if (dist_sq >= ORC_GRAD_DIST_SQ_THRESHOLD) {
                float dist_sq_b = powf(dist_sq, p->b);
                float denom = one + p->a * dist_sq_b;
                float grad_coeff = two_a_b * dist_sq_b / (dist_sq * denom);
                for (int d = 0; d < n_dim; d++) {
                    float delta = yo[d] - ys[d];
                    ADD_TERM(d, two * grad_coeff * delta);
                }
                if (dens) {
                    float phi = one / denom;
                    float dphi_term = p->a * p->b * dist_sq_b / dist_sq * phi;
                    float common = one - p->b * (one - phi);
                    float dr_self =
                        (phi / dens_phi_sum[node]) * (common / expf(dens_re[node]) + dphi_term);
                    float dr_other =
                        (phi / dens_phi_sum[other]) * (common / expf(dens_re[other]) + dphi_term);
                    float cor = dens_coeff *
                                (dens_weight[node] * dr_self + dens_weight[other] * dr_other) /
                                mu[e];
                    float two_cor = two * cor;
                    for (int d = 0; d < n_dim; d++) {
                        float delta = ys[d] - yo[d];
                        float gg = two_cor * delta;
                        if (gg > clip) gg = clip;
                        if (gg < -clip) gg = -clip;
                        ADD_TERM(d, two * gg);
                    }
                }
            }
            for (int32_t s = 0; s < neg_rate; s++) {
                uint32_t k = orc_hash_neg(seed, (uint32_t)node, (uint32_t)epoch, edge_local,
                                          (uint32_t)s, (uint32_t)n);
                if ((int64_t)k == node) continue;
                const float *yk = embd + (int64_t)k * n_dim;
                float dsk = 0.0f;
                for (int d = 0; d < n_dim; d++) {
                    float diff = ys[d] - yk[d];
                    dsk += diff * diff;
                }
                float safe = dsk + rep_eps;
                float dsb = powf(safe, p->b);
                float denom = safe * (one + p->a * dsb);
                float gc = two_gamma_b / denom;
                if (gc > clip) gc = clip;
                if (gc < -clip) gc = -clip;
                for (int d = 0; d < n_dim; d++) {
                    float delta = ys[d] - yk[d];
                    ADD_TERM(d, gc * delta);
                }
            }
        }
#undef ADD_TERM
        if (exact)
            for (int d = 0; d < n_dim; d++) g[d] = (float)gx[d];
        if (has_update) has_update[node] = active;
    }
}

void orc_grad_accum(const float *embd, int64_t n, int32_t n_dim, const uint32_t *off,
                    const uint32_t *csr_edge_idx, const uint32_t *csr_other,
                    const float *next_sample, int32_t epoch, uint32_t seed, int32_t neg_rate,
                    const orc_optim_params *p, const float *mu, const float *dens_re,
                    const float *dens_phi_sum, const float *dens_weight, float dens_coeff,
                    float *grad, uint32_t *has_update, float *abs_out) {
    grad_accum_impl(embd, n, n_dim, off, csr_edge_idx, csr_other, next_sample, epoch, seed,
                    neg_rate, p, mu, dens_re, dens_phi_sum, dens_weight, dens_coeff, grad,
                    has_update, abs_out, 0);
}

void orc_grad_accum_exact(const float *embd, int64_t n, int32_t n_dim, const uint32_t *off,
                          const uint32_t *csr_edge_idx, const uint32_t *csr_other,
                          const float *next_sample, int32_t epoch, uint32_t seed,
                          int32_t neg_rate, const orc_optim_params *p, const float *mu,
                          const float *dens_re, const float *dens_phi_sum,
                          const float *dens_weight, float dens_coeff, float *grad,
                          uint32_t *has_update, float *abs_out) {
    grad_accum_impl(embd, n, n_dim, off, csr_edge_idx, csr_other, next_sample, epoch, seed,
                    neg_rate, p, mu, dens_re, dens_phi_sum, dens_weight, dens_coeff, grad,
                    has_update, abs_out, 1);
}

/* ------------------------------------------------------------------- Adam */

/* launch_adam_update scalars, umap_optimiser_gpu.rs:1168-1175 */
void orc_adam_scalars(const orc_optim_params *p, int32_t epoch, int32_t n_epochs,
                      float *fused_lr, float *epsc) {
    const float one = 1.0f;
    float lr_alpha = p->lr * (one - (float)epoch / (float)n_epochs);
    float t = (float)(epoch + 1);
    float beta1t = powf(p->beta1, t);
    float beta2t = powf(p->beta2, t);
    float sqrt_b2t1 = sqrtf(one - beta2t);
    float ad_scale = sqrt_b2t1 / (one - beta1t);
    *epsc = sqrt_b2t1 * p->eps;
    *fused_lr = lr_alpha * ad_scale;
}

/* cpu_adam_update, umap_optimiser_gpu.rs:1584-1627 */
void orc_adam_update(const float *grad, const uint32_t *has_update, float *m, float *v,
                     float *embd, int64_t n, int32_t n_dim, const orc_optim_params *p,
                     int32_t epoch, int32_t n_epochs) {
    float fused, epsc;
    orc_adam_scalars(p, epoch, n_epochs, &fused, &epsc);
    const float omb1 = 1.0f - p->beta1, omb2 = 1.0f - p->beta2;
    for (int64_t node = 0; node < n; node++) {
        if (!has_update[node]) continue;
        for (int d = 0; d < n_dim; d++) {
            int64_t i = node * n_dim + d;
            float g = grad[i];
            float m_new = m[i] + omb1 * (g - m[i]);
            m[i] = m_new;
            float v_new = v[i] + omb2 * (g * g - v[i]);
            v[i] = v_new;
            embd[i] += fused * m_new / (sqrtf(v_new) + epsc);
        }
    }
}

/* umap_edge_schedule_update, umap_optimiser_gpu.rs:940-955 */
void orc_schedule_update(const float *per, float *next, int64_t n_edges, int32_t epoch) {
    const float et = (float)epoch;
    for (int64_t e = 0; e < n_edges; e++)
        if (next[e] <= et) next[e] += per[e];
}

/* ---------------------------------------------------------------- densMAP */

/* cpu_dens_radii, umap_optimiser_gpu.rs:1409-1453 */
void orc_dens_radii(const float *embd, int64_t n, int32_t n_dim, const uint32_t *off,
                    const uint32_t *csr_other, float a, float b, float *re, float *phi_sum) {
    const float dens_eps = (float)ORC_DENS_EPS;
    for (int64_t node = 0; node < n; node++) {
        const float *ys = embd + node * n_dim;
        float sum_sq = 0.0f, sum_phi = 0.0f;
        for (uint32_t pos = off[node]; pos < off[node + 1]; pos++) {
            const float *yo = embd + (int64_t)csr_other[pos] * n_dim;
            float dist_sq = 0.0f;
            for (int d = 0; d < n_dim; d++) {
                float diff = ys[d] - yo[d];
                dist_sq += diff * diff;
            }
            float dsb = powf(dist_sq, b);
            float phi = 1.0f / (1.0f + a * dsb);
            sum_sq += phi * dist_sq;
            sum_phi += phi;
        }
        float den = sum_phi < FLT_EPSILON ? FLT_EPSILON : sum_phi;
        re[node] = logf(dens_eps + sum_sq / den);
        phi_sum[node] = sum_phi;
    }
}

/* correlation_stats, src/utils/density.rs:464-508 */
void orc_correlation_stats(const float *re, const float *r, int64_t n, float var_shift,
                           double out[3]) {
    const int64_t C = ORC_DENS_REDUCE_CHUNK;
    double denom = (double)((n - 1) > 1 ? (n - 1) : 1);
    double total = 0.0;
    for (int64_t c0 = 0; c0 < n; c0 += C) {
        double s = 0.0;
        int64_t c1 = c0 + C < n ? c0 + C : n;
        for (int64_t i = c0; i < c1; i++) s += (double)re[i];
        total += s;
    }
    double mean = total / (double)(n > 1 ? n : 1);
    double ss = 0.0, cross = 0.0;
    for (int64_t c0 = 0; c0 < n; c0 += C) {
        double s1 = 0.0, s2 = 0.0;
        int64_t c1 = c0 + C < n ? c0 + C : n;
        for (int64_t i = c0; i < c1; i++) {
            double d = (double)re[i] - mean;
            s1 += d * d;
            s2 += d * (double)r[i];
        }
        ss += s1;
        cross += s2;
    }
    double var = ss / denom, cov = cross / denom;
    out[0] = mean;
    out[1] = sqrt(var + (double)var_shift);
    out[2] = cov;
}

/* launch_dens_update host part, umap_optimiser_gpu.rs:1122-1142, and
 * update_density_weights, umap_optimisers.rs:519-537 */
float orc_dens_weights(const float *r, const float *re, int64_t n, float var_shift,
                       float lambda, float mu_tot, float *weight) {
    double st[3];
    orc_correlation_stats(re, r, n, var_shift, st);
    float mean = (float)st[0], sd = (float)st[1], cov = (float)st[2];
    float inv_var = 1.0f / (sd * sd);
    for (int64_t i = 0; i < n; i++) weight[i] = r[i] - cov * (re[i] - mean) * inv_var;
    float denom = (float)((n - 1) > 1 ? (n - 1) : 1);
    return lambda * mu_tot / (sd * denom);
}

/* DensState::is_active, src/utils/density.rs:244-250 */
int orc_dens_is_active(float lambda, float frac, int32_t epoch, int32_t n_epochs) {
    if (lambda <= 0.0f || n_epochs == 0) return 0;
    float progress = (float)(epoch + 1) / (float)n_epochs;
    return progress > 1.0f - frac;
}

/* ------------------------------------------------------ GPU driver replay */

/* optimise_embedding_adam_gpu, umap_optimiser_gpu.rs:1302-1369 */
int orc_optimise_adam_gpu(float *embd, int64_t n, int32_t n_dim, const int64_t *adj_off,
                          const int32_t *adj_cols, const float *adj_vals,
                          const orc_optim_params *p, const float *dens_r, float dens_lambda,
                          float dens_frac, float dens_var_shift, float dens_mu_tot,
                          uint64_t seed) {
    if (n == 0) return 1;
    int64_t ne = 0;
    int st = orc_csr_from_graph(n, adj_off, adj_cols, adj_vals, &ne, NULL, NULL, NULL, NULL,
                                NULL, NULL, NULL);
    if (st) return st;
    uint32_t *off = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    uint32_t *eidx = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(2 * ne));
    uint32_t *oth = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(2 * ne));
    float *per = (float *)malloc(sizeof(float) * (size_t)ne);
    float *next = (float *)malloc(sizeof(float) * (size_t)ne);
    float *mu = (float *)malloc(sizeof(float) * (size_t)ne);
    orc_csr_from_graph(n, adj_off, adj_cols, adj_vals, &ne, off, eidx, oth, per, mu, NULL, NULL);
    memcpy(next, per, sizeof(float) * (size_t)ne);
    size_t nd = (size_t)n * (size_t)n_dim;
    float *m = (float *)calloc(nd, sizeof(float));
    float *v = (float *)calloc(nd, sizeof(float));
    float *g = (float *)calloc(nd, sizeof(float));
    uint32_t *hu = (uint32_t *)calloc((size_t)n, sizeof(uint32_t));
    float *re = NULL, *phi = NULL, *w = NULL;
    if (dens_r) {
        re = (float *)calloc((size_t)n, sizeof(float));
        phi = (float *)calloc((size_t)n, sizeof(float));
        w = (float *)calloc((size_t)n, sizeof(float));
    }
    uint32_t seed32 = (uint32_t)seed;
    for (int32_t e = 0; e < p->n_epochs; e++) {
        int dens_on = dens_r && orc_dens_is_active(dens_lambda, dens_frac, e, p->n_epochs);
        float coeff = 0.0f;
        if (dens_on) {
            orc_dens_radii(embd, n, n_dim, off, oth, p->a, p->b, re, phi);
            coeff = orc_dens_weights(dens_r, re, n, dens_var_shift, dens_lambda, dens_mu_tot, w);
        }
        orc_grad_accum(embd, n, n_dim, off, eidx, oth, next, e, seed32, p->neg_sample_rate, p,
                       mu, dens_on ? re : NULL, phi, w, coeff, g, hu, NULL);
        orc_adam_update(g, hu, m, v, embd, n, n_dim, p, e, p->n_epochs);
        orc_schedule_update(per, next, ne, e);
    }
    free(off); free(eidx); free(oth); free(per); free(next); free(mu);
    free(m); free(v); free(g); free(hu); free(re); free(phi); free(w);
    return 0;
}

/* ------------------------------------------------------------ graph build */

/* local_connectivity_distance, src/data/graph.rs:107-125 */
static float local_connectivity_distance(const float *d, int32_t len, float lc) {
    if (lc <= 0.0f || len == 0) return 0.0f;
    float offset = lc - 1.0f;
    if (offset < 0.0f) offset = 0.0f;
    float whole = floorf(offset);
    float fraction = offset - whole;
    int64_t idx = (int64_t)whole;
    if (idx > len - 1) idx = len - 1;
    if (fraction > 0.0f && idx + 1 < len) return d[idx] * (1.0f - fraction) + d[idx + 1] * fraction;
    return d[idx];
}

/* smooth_knn_dist, src/data/graph.rs:499-565 */
void orc_smooth_knn_dist(const float *dist, int64_t n, int32_t k, float local_connectivity,
                         float bandwidth, int32_t n_iter, float *sigma, float *rho) {
    const double target = log((double)k);
    const double tol = (double)bandwidth;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        const float *d = dist + i * k;
        float adj[1024];
        float rh = local_connectivity_distance(d, k, local_connectivity);
        for (int j = 0; j < k; j++) {
            float a = d[j] - rh;
            adj[j] = a > 0.0f ? a : 0.0f;
        }
        float lo = 0.0f, hi = FLT_MAX, mid = 1.0f;
        for (int it = 0; it < n_iter; it++) {
            float inv_mid = 1.0f / mid;
            float s = 0.0f;
            for (int j = 0; j < k; j++) s += expf(-adj[j] * inv_mid);
            double val = (double)s;
            if (fabs(val - target) < tol) break;
            if (val > target) {
                hi = mid;
                mid = (lo + hi) / 2.0f;
            } else {
                lo = mid;
                if (hi == FLT_MAX) mid *= 2.0f;
                else mid = (lo + hi) / 2.0f;
            }
        }
        sigma[i] = mid;
        rho[i] = rh;
    }
}

/* knn_to_coo, src/data/graph.rs:585-662 */
int64_t orc_knn_to_coo(const int32_t *idx, const float *dist, int64_t n, int32_t k,
                       const float *sigma, const float *rho, int32_t *rows, int32_t *cols,
                       float *vals) {
    int64_t w = 0;
    for (int64_t i = 0; i < n; i++) {
        float s = sigma[i], r = rho[i];
        int smooth = s > 0.0f;
        float inv_sigma = smooth ? 1.0f / s : 0.0f;
        for (int j = 0; j < k; j++) {
            int32_t c = idx[i * k + j];
            if (c == i) continue;
            float a = dist[i * k + j] - r;
            if (a < 0.0f) a = 0.0f;
            float wt = smooth ? expf(-a * inv_sigma) : (a > 0.0f ? 0.0f : 1.0f);
            rows[w] = (int32_t)i;
            cols[w] = c;
            vals[w] = wt;
            w++;
        }
    }
    return w;
}

typedef struct {
    int32_t c;
    float w;
    int64_t ord;
} orc_entry;

static int cmp_entry(const void *a, const void *b) {
    const orc_entry *x = (const orc_entry *)a, *y = (const orc_entry *)b;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    return x->ord < y->ord ? -1 : (x->ord > y->ord);
}

/* symmetrise_csr + symmetrise_graph, src/data/graph.rs:185-468, 684-694 */
int64_t orc_symmetrise(int64_t n, const int32_t *rows, const int32_t *cols, const float *vals,
                       int64_t nnz, float mix_weight, int32_t *out_rows, int32_t *out_cols,
                       float *out_vals) {
    if (n == 0 || nnz == 0) return 0;
    const float directed = 1.0f - mix_weight;
    /* coo_to_sorted_csr: counting sort by row, then sort each row by column */
    int64_t *ip = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t e = 0; e < nnz; e++) ip[rows[e] + 1]++;
    for (int64_t i = 0; i < n; i++) ip[i + 1] += ip[i];
    orc_entry *ent = (orc_entry *)malloc(sizeof(orc_entry) * (size_t)nnz);
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    memcpy(cur, ip, sizeof(int64_t) * (size_t)n);
    for (int64_t e = 0; e < nnz; e++) {
        int64_t p = cur[rows[e]]++;
        ent[p].c = cols[e];
        ent[p].w = vals[e];
        ent[p].ord = e;
    }
    for (int64_t i = 0; i < n; i++)
        qsort(ent + ip[i], (size_t)(ip[i + 1] - ip[i]), sizeof(orc_entry), cmp_entry);
    /* transpose_csr: counting sort visiting rows in order */
    int64_t *tp = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t e = 0; e < nnz; e++) tp[ent[e].c + 1]++;
    for (int64_t i = 0; i < n; i++) tp[i + 1] += tp[i];
    orc_entry *tent = (orc_entry *)malloc(sizeof(orc_entry) * (size_t)nnz);
    memcpy(cur, tp, sizeof(int64_t) * (size_t)n);
    for (int64_t r = 0; r < n; r++)
        for (int64_t p = ip[r]; p < ip[r + 1]; p++) {
            int64_t q = cur[ent[p].c]++;
            tent[q].c = (int32_t)r;
            tent[q].w = ent[p].w;
            tent[q].ord = 0;
        }
    /* merge_row for every row, in row order */
    int64_t written = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t ia = ip[i], ea = ip[i + 1], ib = tp[i], eb = tp[i + 1];
        while (ia < ea || ib < eb) {
            int32_t col;
            if (ia < ea && ib < eb) col = ent[ia].c < tent[ib].c ? ent[ia].c : tent[ib].c;
            else if (ia < ea) col = ent[ia].c;
            else col = tent[ib].c;
            float wij = 0.0f, wji = 0.0f;
            while (ia < ea && ent[ia].c == col) wij = ent[ia++].w;
            while (ib < eb && tent[ib].c == col) wji = tent[ib++].w;
            float uni = wij + wji - wij * wji;
            float wsym = mix_weight * uni + directed * wij;
            if (wsym > 0.0f) {
                out_rows[written] = (int32_t)i;
                out_cols[written] = col;
                out_vals[written] = wsym;
                written++;
            }
        }
    }
    free(ip); free(ent); free(cur); free(tp); free(tent);
    return written;
}

/* filter_weak_edges, src/data/graph.rs:776-853 */
int64_t orc_filter_weak_edges(int32_t *rows, int32_t *cols, float *vals, int64_t nnz,
                              int32_t n_epochs) {
    float max_w = 0.0f;
    for (int64_t e = 0; e < nnz; e++)
        if (vals[e] > max_w) max_w = vals[e];
    float thr = max_w / (float)n_epochs;
    int64_t w = 0;
    for (int64_t e = 0; e < nnz; e++)
        if (vals[e] >= thr) {
            rows[w] = rows[e];
            cols[w] = cols[e];
            vals[w] = vals[e];
            w++;
        }
    return w;
}

/* ------------------------------------------------------------------- kNN */

#define KNN_MAXK 256
#define KNN_QT 8 /* queries per tile: each candidate row is read once per 8 queries */

/* Exhaustive kNN of the queries qrows[0..nq) (nq <= KNN_QT) against all n rows:
 * squared distances summed over dims in order (fmaf for f32, plain in f64), the
 * k smallest kept in an insertion list where an equal distance never displaces
 * an earlier (smaller) index.  Per (query, candidate) pair the operations are
 * exactly those of a scalar double loop; the tile only interleaves 8 independent
 * pairs so the sums vectorise across queries. */
#define KNN_TILE(NAME, T, ACC)                                                               \
    static void NAME(const float *x, int64_t n, int32_t d, int32_t k, const int64_t *qrows,   \
                     int nq, int32_t *idx, T *dist) {                                          \
        T *qt = (T *)calloc((size_t)d * KNN_QT, sizeof(T));                                   \
        T bd[KNN_QT][KNN_MAXK];                                                                \
        int32_t bi[KNN_QT][KNN_MAXK];                                                          \
        int cnt[KNN_QT] = {0};                                                                 \
        for (int q = 0; q < nq; q++)                                                           \
            for (int t = 0; t < d; t++) qt[(size_t)t * KNN_QT + q] = (T)x[qrows[q] * d + t];   \
        for (int64_t j = 0; j < n; j++) {                                                      \
            const float *xj = x + j * d;                                                       \
            T s[KNN_QT];                                                                       \
            for (int q = 0; q < KNN_QT; q++) s[q] = 0;                                         \
            for (int t = 0; t < d; t++) {                                                      \
                const T xt = (T)xj[t];                                                         \
                const T *qrow = qt + (size_t)t * KNN_QT;                                       \
                for (int q = 0; q < KNN_QT; q++) {                                             \
                    T df = qrow[q] - xt;                                                       \
                    s[q] = ACC(df, s[q]);                                                      \
                }                                                                              \
            }                                                                                  \
            for (int q = 0; q < nq; q++) {                                                     \
                if (j == qrows[q]) continue;                                                   \
                if (cnt[q] == k && !(s[q] < bd[q][k - 1])) continue;                           \
                int p = cnt[q] < k ? cnt[q] : k - 1;                                           \
                while (p > 0 && (s[q] < bd[q][p - 1])) {                                       \
                    bd[q][p] = bd[q][p - 1];                                                   \
                    bi[q][p] = bi[q][p - 1];                                                   \
                    p--;                                                                       \
                }                                                                              \
                bd[q][p] = s[q];                                                               \
                bi[q][p] = (int32_t)j;                                                         \
                if (cnt[q] < k) cnt[q]++;                                                      \
            }                                                                                  \
        }                                                                                      \
        for (int q = 0; q < nq; q++)                                                           \
            for (int t = 0; t < k; t++) {                                                      \
                idx[(int64_t)q * k + t] = t < cnt[q] ? bi[q][t] : -1;                          \
                dist[(int64_t)q * k + t] = t < cnt[q] ? bd[q][t] : (T)INFINITY;                \
            }                                                                                  \
        free(qt);                                                                              \
    }

#define ACC_F32(df, s) fmaf((df), (df), (s))
#define ACC_F64(df, s) ((s) + (df) * (df))
KNN_TILE(knn_tile_f32, float, ACC_F32)
KNN_TILE(knn_tile_f64, double, ACC_F64)

void orc_knn_exact_f32(const float *x, int64_t n, int32_t d, int32_t k, int64_t q0, int64_t q1,
                       int32_t *idx, float *dist) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t a = q0; a < q1; a += KNN_QT) {
        int64_t rows[KNN_QT];
        int nq = (int)(q1 - a < KNN_QT ? q1 - a : KNN_QT);
        for (int q = 0; q < nq; q++) rows[q] = a + q;
        knn_tile_f32(x, n, d, k, rows, nq, idx + (a - q0) * k, dist + (a - q0) * k);
    }
}

void orc_knn_exact_f64(const float *x, int64_t n, int32_t d, int32_t k, int64_t q0, int64_t q1,
                       int32_t *idx, double *dist) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t a = q0; a < q1; a += KNN_QT) {
        int64_t rows[KNN_QT];
        int nq = (int)(q1 - a < KNN_QT ? q1 - a : KNN_QT);
        for (int q = 0; q < nq; q++) rows[q] = a + q;
        knn_tile_f64(x, n, d, k, rows, nq, idx + (a - q0) * k, dist + (a - q0) * k);
    }
}

void orc_knn_exact_f64_rows(const float *x, int64_t n, int32_t d, int32_t k,
                            const int64_t *rows, int64_t n_rows, int32_t *idx, double *dist) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t a = 0; a < n_rows; a += KNN_QT) {
        int nq = (int)(n_rows - a < KNN_QT ? n_rows - a : KNN_QT);
        knn_tile_f64(x, n, d, k, rows + a, nq, idx + a * k, dist + a * k);
    }
}

/* --------------------------------------------------------------- DensState */

/* DensState::new, src/utils/density.rs:181-201, with edge_distances_from_knn
 * :309-334, original_log_radii :358-398 and standardise :410-438 */
int orc_dens_state(int64_t n, const int32_t *rows, const int32_t *cols, const float *vals,
                   int64_t nnz, const int32_t *knn_idx, const float *knn_dist, int32_t k,
                   int32_t distances_squared, float *r_out, float *mu_sum_out,
                   float *mu_tot_out) {
    if (n == 0) return 1;
    double *ro = (double *)calloc((size_t)n, sizeof(double));
    double *ms = (double *)calloc((size_t)n, sizeof(double));
    for (int64_t e = 0; e < nnz; e++) {
        int32_t i = rows[e], j = cols[e];
        float raw = 0.0f;
        int found = 0;
        for (int t = 0; t < k; t++)
            if (knn_idx[(int64_t)i * k + t] == j) {
                raw = knn_dist[(int64_t)i * k + t];
                found = 1;
                break;
            }
        if (!found)
            for (int t = 0; t < k; t++)
                if (knn_idx[(int64_t)j * k + t] == i) {
                    raw = knn_dist[(int64_t)j * k + t];
                    break;
                }
        float dsq = distances_squared ? raw : raw * raw;
        double mu = (double)vals[e];
        ro[i] += mu * (double)dsq;
        ms[i] += mu;
    }
    for (int64_t i = 0; i < n; i++) {
        double den = ms[i] > DBL_EPSILON ? ms[i] : DBL_EPSILON;
        ro[i] = log(ORC_DENS_EPS + ro[i] / den);
    }
    int status = 0;
    if (n < 2) {
        status = 3;
    } else {
        double mean = 0.0;
        for (int64_t i = 0; i < n; i++) mean += ro[i];
        mean /= (double)n;
        double var = 0.0;
        for (int64_t i = 0; i < n; i++) {
            double dd = ro[i] - mean;
            var += dd * dd;
        }
        var /= (double)(n - 1);
        double sd = sqrt(var);
        if (isnan(sd) || sd <= DBL_EPSILON) status = 3;
        else
            for (int64_t i = 0; i < n; i++) ro[i] = (ro[i] - mean) / sd;
    }
    if (status == 0) {
        double tot = 0.0;
        for (int64_t i = 0; i < n; i++) {
            r_out[i] = (float)ro[i];
            mu_sum_out[i] = (float)ms[i];
            tot += (double)mu_sum_out[i];
        }
        *mu_tot_out = (float)tot;
    }
    free(ro);
    free(ms);
    return status;
}

/* -------------------------------------------------------------- fit_params */

/* the 300 (x, target) points of fit_params, src/training/umap_optimisers.rs:145-164 */
#define ORC_FIT_POINTS 300
static void fit_points(float min_dist, float spread, float *xv, float *yv) {
    float max_x = spread * 3.0f;
    float step = max_x / (float)(ORC_FIT_POINTS - 1);
    for (int i = 0; i < ORC_FIT_POINTS; i++) {
        float x = step * (float)i;
        xv[i] = x;
        yv[i] = x < min_dist ? 1.0f : expf(-(x - min_dist) / spread);
    }
}

/* one iteration's (grad_a, grad_b), :170-193 */
static void fit_grad(const float *xv, const float *yv, float a, float b, float *ga_out,
                     float *gb_out) {
    const float two = 2.0f;
    float ga = 0.0f, gb = 0.0f;
    for (int i = 0; i < ORC_FIT_POINTS; i++) {
        float x = xv[i];
        if (x <= 0.0f) continue;
        float x2b = powf(x, two * b);
        float denom = 1.0f + a * x2b;
        float pred = 1.0f / denom;
        float err = pred - yv[i];
        ga += err * (-x2b / (denom * denom));
        float lx = logf(x);
        gb += err * (-two * a * x2b * lx / (denom * denom));
    }
    *ga_out = ga / (float)ORC_FIT_POINTS;
    *gb_out = gb / (float)ORC_FIT_POINTS;
}

void orc_fit_params_grad(float min_dist, float spread, float a, float b, float *ga, float *gb) {
    float xv[ORC_FIT_POINTS], yv[ORC_FIT_POINTS];
    fit_points(min_dist, spread, xv, yv);
    fit_grad(xv, yv, a, b, ga, gb);
}

/* UmapOptimParams::fit_params, src/training/umap_optimisers.rs:139-209 */
void orc_fit_params(float min_dist, float spread, int32_t n_iter, float *a_out, float *b_out) {
    if (n_iter <= 0) n_iter = 300;
    float xv[ORC_FIT_POINTS], yv[ORC_FIT_POINTS];
    fit_points(min_dist, spread, xv, yv);
    float a = 1.0f, b = 1.0f;
    for (int it = 0; it < n_iter; it++) {
        float ga, gb;
        fit_grad(xv, yv, a, b, &ga, &gb);
        a -= 1.0f * ga;
        b -= 1.0f * gb;
        a = a < 0.001f ? 0.001f : a;
        a = a > 10.0f ? 10.0f : a;
        b = b < 0.1f ? 0.1f : b;
        b = b > 2.0f ? 2.0f : b;
    }
    *a_out = a;
    *b_out = b;
}

/* ------------------------------------------------- CPU optimisers (baseline) */

typedef struct {
    float b, max_val, inv_step;
    float *table;
    int size;
} orc_lut;

/* FastPowLut::new / get, src/training/umap_optimisers.rs:325-397 */
static void lut_init(orc_lut *L, float b, float max_val, int size) {
    L->b = b;
    L->max_val = max_val;
    L->size = size;
    float step = max_val / (float)(size - 1);
    L->inv_step = 1.0f / step;
    L->table = (float *)malloc(sizeof(float) * (size_t)size);
    for (int i = 0; i < size; i++) L->table[i] = powf(step * (float)i, b);
}

static inline float lut_get(const orc_lut *L, float x) {
    if (x >= L->max_val) return powf(x, L->b);
    float idx_f = x * L->inv_step;
    int64_t idx = idx_f > 0.0f ? (int64_t)idx_f : 0;
    if (idx >= L->size - 1) return L->table[L->size - 1];
    float rem = idx_f - (float)idx;
    float y0 = L->table[idx], y1 = L->table[idx + 1];
    return y0 + rem * (y1 - y0);
}

/* the parallel optimiser's LUT (65 536 entries on [0, 25]) applied to x[] */
void orc_lut_pow(float b, const float *x, int64_t n, float *out) {
    orc_lut L;
    lut_init(&L, b, 25.0f, 65536);
    for (int64_t i = 0; i < n; i++) out[i] = lut_get(&L, x[i]);
    free(L.table);
}

/* density_edge_coeff, src/training/umap_optimisers.rs:564-592, with its
 * max(T::epsilon()) floors on phi_sum and mu */
static inline float dens_edge_coeff(float a, float b, float dist_sq, float dist_sq_b,
                                    float denom, float phi_i, float phi_o, float re_i,
                                    float re_o, float w_i, float w_o, float coeff, float mu_e) {
    const float one = 1.0f, feps = FLT_EPSILON;
    float phi = one / denom;
    float dphi_term = a * b * dist_sq_b / dist_sq * phi;
    float common = one - b * (one - phi);
    float pi_ = phi_i > feps ? phi_i : feps;
    float po_ = phi_o > feps ? phi_o : feps;
    float dr_self = (phi / pi_) * (common / expf(re_i) + dphi_term);
    float dr_other = (phi / po_) * (common / expf(re_o) + dphi_term);
    float mue = mu_e > feps ? mu_e : feps;
    return coeff * (w_i * dr_self + w_o * dr_other) / mue;
}

float orc_density_edge_coeff(float a, float b, float dist_sq, float dist_sq_b, float denom,
                             float phi_sum_i, float phi_sum_o, float re_i, float re_o, float w_i,
                             float w_o, float coeff, float mu_edge) {
    return dens_edge_coeff(a, b, dist_sq, dist_sq_b, denom, phi_sum_i, phi_sum_o, re_i, re_o, w_i,
                           w_o, coeff, mu_edge);
}

/* optimise_embedding_adam_parallel, src/training/umap_optimisers.rs:1170-1492,
 * with the densMAP pass 0 (:1303-1317: accumulate_umap_radii :468-511,
 * update_density_weights :519-537) and density_edge_coeff (:564-592, with its
 * max(eps) floors) when dens_r is given. */
int orc_optimise_adam_parallel_dens(float *embd, int64_t n, int32_t n_dim,
                                    const int64_t *adj_off, const int32_t *adj_cols,
                                    const float *adj_vals, const orc_optim_params *p,
                                    uint64_t seed, int32_t use_lut, int32_t n_threads,
                                    const float *dens_r, float dens_lambda, float dens_frac,
                                    float dens_var_shift, float dens_mu_tot) {
    if (n == 0) return 1;
    int64_t ne = 0;
    int st = orc_csr_from_graph(n, adj_off, adj_cols, adj_vals, &ne, NULL, NULL, NULL, NULL,
                                NULL, NULL, NULL);
    if (st) return st;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    uint32_t *off = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    uint32_t *eidx = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(2 * ne));
    uint32_t *oth = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(2 * ne));
    float *per = (float *)malloc(sizeof(float) * (size_t)ne);
    float *mu = (float *)malloc(sizeof(float) * (size_t)ne);
    int32_t *ei = (int32_t *)malloc(sizeof(int32_t) * (size_t)ne);
    orc_csr_from_graph(n, adj_off, adj_cols, adj_vals, &ne, off, eidx, oth, per, mu, ei, NULL);
    float *next = (float *)malloc(sizeof(float) * (size_t)ne);
    float *per_neg = (float *)malloc(sizeof(float) * (size_t)ne);
    float *next_neg = (float *)malloc(sizeof(float) * (size_t)ne);
    const float nsr = (float)p->neg_sample_rate;
    for (int64_t e = 0; e < ne; e++) {
        next[e] = per[e];
        per_neg[e] = per[e] / nsr;
        next_neg[e] = per_neg[e];
    }
    size_t nd = (size_t)n * (size_t)n_dim;
    float *m = (float *)calloc(nd, sizeof(float));
    float *v = (float *)calloc(nd, sizeof(float));
    float *g = (float *)calloc(nd, sizeof(float));
    unsigned char *hu = (unsigned char *)calloc((size_t)n, 1);
    const float two = 2.0f;
    const float two_a_b = two * p->a * p->b, two_gamma_b = two * p->gamma * p->b;
    const float clip = 4.0f, ceps = 0.001f;
    const int b_is_one = fabsf(p->b - 1.0f) < 1e-10f;
    orc_lut L;
    lut_init(&L, p->b, 25.0f, 65536);
    const uint32_t seed32 = (uint32_t)seed;
    const float nef = (float)p->n_epochs;
    float *re = NULL, *phis = NULL, *dw = NULL;
    if (dens_r) {
        re = (float *)calloc((size_t)n, sizeof(float));
        phis = (float *)calloc((size_t)n, sizeof(float));
        dw = (float *)calloc((size_t)n, sizeof(float));
    }
    const float feps = FLT_EPSILON;
    for (int32_t epoch = 0; epoch < p->n_epochs; epoch++) {
        const float lr = p->lr * (1.0f - (float)epoch / nef);
        const float et = (float)epoch;
        const float t = (float)(epoch + 1);
        const float sqrt_b2t1 = sqrtf(1.0f - powf(p->beta2, t));
        const float ad_scale = sqrt_b2t1 / (1.0f - powf(p->beta1, t));
        const float epsc = sqrt_b2t1 * p->eps;
        const int dens_on = dens_r && orc_dens_is_active(dens_lambda, dens_frac, epoch, p->n_epochs);
        float dcoeff = 0.0f;
        if (dens_on) {
#pragma omp parallel for schedule(dynamic, 256)
            for (int64_t i = 0; i < n; i++) {
                const float *yi = embd + i * n_dim;
                float ssq = 0.0f, sphi = 0.0f;
                for (uint32_t pos = off[i]; pos < off[i + 1]; pos++) {
                    const float *yo = embd + (int64_t)oth[pos] * n_dim;
                    float dsq = 0.0f;
                    for (int d = 0; d < n_dim; d++) {
                        float df = yi[d] - yo[d];
                        dsq += df * df;
                    }
                    float dsb = b_is_one ? dsq : lut_get(&L, dsq);
                    float phi = 1.0f / (1.0f + p->a * dsb);
                    ssq += phi * dsq;
                    sphi += phi;
                }
                phis[i] = sphi;
                re[i] = logf((float)ORC_DENS_EPS + ssq / (sphi > feps ? sphi : feps));
            }
            dcoeff = orc_dens_weights(dens_r, re, n, dens_var_shift, dens_lambda, dens_mu_tot, dw);
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; i++) {
            float *gi = g + i * n_dim;
            for (int d = 0; d < n_dim; d++) gi[d] = 0.0f;
            const float *yi = embd + i * n_dim;
            int any = 0;
            uint32_t slot = 0;
            for (uint32_t pos = off[i]; pos < off[i + 1]; pos++, slot++) {
                uint32_t e = eidx[pos];
                if (next[e] > et) continue;
                any = 1;
                uint32_t o = oth[pos];
                const float *yo = embd + (int64_t)o * n_dim;
                float dsq = 0.0f;
                for (int d = 0; d < n_dim; d++) {
                    float df = yi[d] - yo[d];
                    dsq += df * df;
                }
                if (dsq >= 1e-8f) {
                    float dsb = b_is_one ? dsq : (use_lut ? lut_get(&L, dsq) : powf(dsq, p->b));
                    float denom = 1.0f + p->a * dsb;
                    float gc = two_a_b * dsb / (dsq * denom);
                    for (int d = 0; d < n_dim; d++) gi[d] += two * gc * (yo[d] - yi[d]);
                    if (dens_on) {
                        float cor = dens_edge_coeff(p->a, p->b, dsq, dsb, denom, phis[i],
                                                    phis[o], re[i], re[o], dw[i], dw[o], dcoeff,
                                                    mu[e]);
                        float two_cor = two * cor;
                        for (int d = 0; d < n_dim; d++) {
                            float gg = two_cor * (yi[d] - yo[d]);
                            if (gg > clip) gg = clip;
                            if (gg < -clip) gg = -clip;
                            gi[d] += two * gg;
                        }
                    }
                }
                if (ei[e] == (int32_t)i) { /* is_smaller */
                    float nf = floorf((et - next_neg[e]) / per_neg[e]);
                    int64_t nneg = nf > 0.0f ? (int64_t)nf : 0;
                    for (int64_t s = 0; s < nneg; s++) {
                        uint32_t kk = orc_hash_neg(seed32, (uint32_t)i, (uint32_t)epoch, slot,
                                                   (uint32_t)s, (uint32_t)n);
                        if ((int64_t)kk == i) continue;
                        const float *yk = embd + (int64_t)kk * n_dim;
                        float dk = 0.0f;
                        for (int d = 0; d < n_dim; d++) {
                            float df = yi[d] - yk[d];
                            dk += df * df;
                        }
                        float safe = dk + ceps;
                        float dsb = b_is_one ? safe : (use_lut ? lut_get(&L, safe) : powf(safe, p->b));
                        float denom = safe * (1.0f + p->a * dsb);
                        float gc = two_gamma_b / denom;
                        if (gc > clip) gc = clip;
                        if (gc < -clip) gc = -clip;
                        for (int d = 0; d < n_dim; d++) gi[d] += gc * (yi[d] - yk[d]);
                    }
                }
            }
            hu[i] = (unsigned char)any;
        }
        const float omb1 = 1.0f - p->beta1, omb2 = 1.0f - p->beta2;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            if (!hu[i]) continue;
            for (int d = 0; d < n_dim; d++) {
                int64_t q = i * n_dim + d;
                float gg = g[q];
                m[q] += omb1 * (gg - m[q]);
                v[q] += omb2 * (gg * gg - v[q]);
                embd[q] += lr * ad_scale * m[q] / (sqrtf(v[q]) + epsc);
            }
        }
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < ne; e++) {
            if (next[e] <= et) {
                next[e] += per[e];
                float nf = floorf((et - next_neg[e]) / per_neg[e]);
                int64_t nneg = nf > 0.0f ? (int64_t)nf : 0;
                next_neg[e] += (float)nneg * per_neg[e];
            }
        }
    }
    free(L.table);
    free(off); free(eidx); free(oth); free(per); free(mu); free(ei); free(next);
    free(per_neg); free(next_neg); free(m); free(v); free(g); free(hu);
    free(re); free(phis); free(dw);
    return 0;
}

int orc_optimise_adam_parallel(float *embd, int64_t n, int32_t n_dim, const int64_t *adj_off,
                               const int32_t *adj_cols, const float *adj_vals,
                               const orc_optim_params *p, uint64_t seed, int32_t use_lut,
                               int32_t n_threads) {
    return orc_optimise_adam_parallel_dens(embd, n, n_dim, adj_off, adj_cols, adj_vals, p, seed,
                                           use_lut, n_threads, NULL, 0.0f, 0.0f, 0.0f, 0.0f);
}

/* optimise_embedding_sgd, src/training/umap_optimisers.rs:653-865 (plain UMAP) */
int orc_optimise_sgd(float *embd, int64_t n, int32_t n_dim, const int64_t *adj_off,
                     const int32_t *adj_cols, const float *adj_vals, const orc_optim_params *p,
                     uint64_t seed) {
    if (n == 0) return 1;
    int64_t ne = adj_off[n];
    if (ne == 0) return 2;
    float max_w = 0.0f;
    for (int64_t e = 0; e < ne; e++)
        if (adj_vals[e] > max_w) max_w = adj_vals[e];
    float *per = (float *)malloc(sizeof(float) * (size_t)ne);
    float *next = (float *)malloc(sizeof(float) * (size_t)ne);
    float *per_neg = (float *)malloc(sizeof(float) * (size_t)ne);
    float *next_neg = (float *)malloc(sizeof(float) * (size_t)ne);
    int32_t *src = (int32_t *)malloc(sizeof(int32_t) * (size_t)ne);
    for (int64_t i = 0; i < n; i++)
        for (int64_t q = adj_off[i]; q < adj_off[i + 1]; q++) src[q] = (int32_t)i;
    const float nsr = (float)p->neg_sample_rate;
    for (int64_t e = 0; e < ne; e++) {
        float norm = adj_vals[e] / max_w;
        per[e] = norm > 0.0f ? 1.0f / norm : 1e8f;
        next[e] = per[e];
        per_neg[e] = per[e] / nsr;
        next_neg[e] = per_neg[e];
    }
    const float two = 2.0f;
    const float two_a_b = two * p->a * p->b, two_gamma_b = two * p->gamma * p->b;
    const float clip = 4.0f;
    const int b_is_one = fabsf(p->b - 1.0f) < 1e-10f;
    const int b_is_half = fabsf(p->b - 0.5f) < 1e-10f;
#define FPOW(x) (b_is_one ? (x) : (b_is_half ? sqrtf(x) : powf((x), p->b)))
    const uint32_t seed32 = (uint32_t)seed;
    const float nef = (float)p->n_epochs;
    for (int32_t epoch = 0; epoch < p->n_epochs; epoch++) {
        const float lr = p->lr * (1.0f - (float)epoch / nef);
        const float et = (float)epoch;
        for (int64_t e = 0; e < ne; e++) {
            if (next[e] > et) continue;
            int64_t i = src[e], j = adj_cols[e];
            float *yi = embd + i * n_dim, *yj = embd + j * n_dim;
            float dsq = 0.0f;
            for (int d = 0; d < n_dim; d++) {
                float df = yi[d] - yj[d];
                dsq += df * df;
            }
            if (dsq >= 1e-8f) {
                float dsb = FPOW(dsq);
                float denom = 1.0f + p->a * dsb;
                float gc = two_a_b * dsb / (dsq * denom);
                for (int d = 0; d < n_dim; d++) {
                    float delta = yj[d] - yi[d];
                    float gd = gc * delta;
                    if (gd > clip) gd = clip;
                    if (gd < -clip) gd = -clip;
                    yi[d] += gd * lr;
                    yj[d] -= gd * lr;
                }
            }
            next[e] += per[e];
            float nf = floorf((et - next_neg[e]) / per_neg[e]);
            int64_t nneg = nf > 0.0f ? (int64_t)nf : 0;
            for (int64_t s = 0; s < nneg; s++) {
                uint32_t kk = orc_hash_neg(seed32, (uint32_t)i, (uint32_t)epoch,
                                           (uint32_t)(e - adj_off[i]), (uint32_t)s, (uint32_t)n);
                if ((int64_t)kk == i) continue;
                float *yk = embd + (int64_t)kk * n_dim;
                float dk = 0.0f;
                for (int d = 0; d < n_dim; d++) {
                    float df = yi[d] - yk[d];
                    dk += df * df;
                }
                float safe = dk + 0.001f;
                float dsb = FPOW(safe);
                float denom = safe * (1.0f + p->a * dsb);
                float gc = two_gamma_b / denom;
                if (gc > clip) gc = clip;
                if (gc < -clip) gc = -clip;
                for (int d = 0; d < n_dim; d++) yi[d] += gc * (yi[d] - yk[d]) * lr;
            }
            next_neg[e] += (float)nneg * per_neg[e];
        }
    }
#undef FPOW
    free(per); free(next); free(per_neg); free(next_neg); free(src);
    return 0;
}

/* ------------------------------------------------------------ spectral init */

/* Counter-based stand-ins for the reference's StdRng draws: its streams
 * (rand 0.9.5) cannot be reproduced here, their distributions can.  stream
 * separates the uses (start vector, noise, random fill). */
static uint64_t smix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double u01(uint64_t seed, uint64_t stream, uint64_t i) {
    return (double)(smix(smix(seed * 0x100000001B3ull + stream) + i) >> 11) *
           (1.0 / 9007199254740992.0);
}
static double std_normal(uint64_t seed, uint64_t stream, uint64_t i) { /* Box-Muller */
    double u = u01(seed, stream, 2 * i), v = u01(seed, stream, 2 * i + 1);
    if (u < 1e-300) u = 1e-300;
    return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

/* Symmetric eigenproblem of a small dense matrix (the tridiagonal T of
 * Lanczos, src/utils/math.rs:366-388, which the reference hands to faer's
 * self_adjoint_eigen): cyclic Jacobi rotations.  A [m][m] is destroyed;
 * eigenvalues ascending, eigenvectors as the columns of V [m][m]. */
static void jacobi_eigen(double *A, int m, double *evals, double *V) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) V[i * m + j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) off += A[p * m + q] * A[p * m + q];
        if (off < 1e-30) break;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) {
                double apq = A[p * m + q];
                if (fabs(apq) < 1e-300) continue;
                double theta = (A[q * m + q] - A[p * m + p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; k++) { /* columns p, q */
                    double akp = A[k * m + p], akq = A[k * m + q];
                    A[k * m + p] = c * akp - s * akq;
                    A[k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; k++) { /* rows p, q */
                    double apk = A[p * m + k], aqk = A[q * m + k];
                    A[p * m + k] = c * apk - s * aqk;
                    A[q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; k++) {
                    double vkp = V[k * m + p], vkq = V[k * m + q];
                    V[k * m + p] = c * vkp - s * vkq;
                    V[k * m + q] = s * vkp + c * vkq;
                }
            }
    }
    /* sort ascending (selection sort keeps equal values in index order) */
    int *ord = (int *)malloc(sizeof(int) * (size_t)m);
    for (int i = 0; i < m; i++) ord[i] = i;
    for (int i = 0; i < m; i++)
        for (int j = i + 1; j < m; j++)
            if (A[ord[j] * m + ord[j]] < A[ord[i] * m + ord[i]]) {
                int t = ord[i];
                ord[i] = ord[j];
                ord[j] = t;
            }
    double *W = (double *)malloc(sizeof(double) * (size_t)m * m);
    memcpy(W, V, sizeof(double) * (size_t)m * m);
    for (int j = 0; j < m; j++) {
        evals[j] = A[ord[j] * m + ord[j]];
        for (int i = 0; i < m; i++) V[i * m + j] = W[i * m + ord[j]];
    }
    free(W);
    free(ord);
}

static double dotp(const double *a, const double *b, int64_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static) if (n > 65536)
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* lanczos_eigendecomp + select_eigenpairs (smallest), src/utils/math.rs:401-627,
 * on the CSR matrix M (f64): min(4 n_eigs + 30, n) iterations with full
 * modified Gram-Schmidt, the tridiagonal problem, Ritz vectors of the n_eigs
 * smallest Ritz values normalised and cast to f32 into evecs [n][n_eigs].
 * Returns how many pairs were produced (fewer on an early breakdown). */
static int lanczos_smallest(int64_t n, const int64_t *ip, const int32_t *ix, const double *dv,
                            int n_eigs, uint64_t seed, float *evecs) {
    int n_iter = n_eigs * 4 + 30;
    if ((int64_t)n_iter > n) n_iter = (int)n;
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    double *v_old = (double *)calloc((size_t)n, sizeof(double));
    double *w = (double *)malloc(sizeof(double) * (size_t)n);
    double *basis = (double *)malloc(sizeof(double) * (size_t)n * n_iter);
    double *alpha = (double *)calloc((size_t)n_iter, sizeof(double));
    double *beta = (double *)calloc((size_t)n_iter, sizeof(double));
    for (int64_t i = 0; i < n; i++) v[i] = u01(seed, 1, (uint64_t)i) - 0.5;
    double nv = sqrt(dotp(v, v, n));
    for (int64_t i = 0; i < n; i++) v[i] /= nv;
    int n_done = 0;
    for (int j = 0; j < n_iter; j++) {
        memcpy(basis + (size_t)j * n, v, sizeof(double) * (size_t)n);
        n_done = j + 1;
#pragma omp parallel for schedule(static) if (n > 65536)
        for (int64_t i = 0; i < n; i++) {
            double acc = 0.0;
            for (int64_t q = ip[i]; q < ip[i + 1]; q++) acc += dv[q] * v[ix[q]];
            w[i] = acc;
        }
        alpha[j] = dotp(w, v, n);
        const double a = alpha[j], bprev = j > 0 ? beta[j - 1] : 0.0;
#pragma omp parallel for schedule(static) if (n > 65536)
        for (int64_t i = 0; i < n; i++) {
            w[i] -= a * v[i];
            if (j > 0) w[i] -= bprev * v_old[i];
        }
        for (int k = 0; k <= j; k++) { /* modified Gram-Schmidt against every basis vector */
            const double *bk = basis + (size_t)k * n;
            const double coeff = dotp(w, bk, n);
#pragma omp parallel for schedule(static) if (n > 65536)
            for (int64_t i = 0; i < n; i++) w[i] -= coeff * bk[i];
        }
        beta[j] = sqrt(dotp(w, w, n));
        if (beta[j] < 1e-12) break;
        memcpy(v_old, v, sizeof(double) * (size_t)n);
        for (int64_t i = 0; i < n; i++) v[i] = w[i] / beta[j];
    }
    const int m = n_done;
    double *T = (double *)calloc((size_t)m * m, sizeof(double));
    double *ev = (double *)malloc(sizeof(double) * (size_t)m);
    double *V = (double *)malloc(sizeof(double) * (size_t)m * m);
    for (int i = 0; i < m; i++) {
        T[i * m + i] = alpha[i];
        if (i + 1 < m) T[i * m + i + 1] = T[(i + 1) * m + i] = beta[i];
    }
    jacobi_eigen(T, m, ev, V);
    const int n_sel = n_eigs < m ? n_eigs : m;
    double *e = (double *)malloc(sizeof(double) * (size_t)n);
    for (int s = 0; s < n_sel; s++) {
        for (int64_t i = 0; i < n; i++) e[i] = 0.0;
        for (int j = 0; j < m; j++) { /* j outer: every element accumulates in ascending j */
            const double cj = V[j * m + s];
            const double *bj = basis + (size_t)j * n;
            for (int64_t i = 0; i < n; i++) e[i] += bj[i] * cj;
        }
        double ne = 0.0;
        for (int64_t i = 0; i < n; i++) ne += e[i] * e[i];
        ne = sqrt(ne);
        for (int64_t i = 0; i < n; i++) evecs[i * n_eigs + s] = (float)(ne > 0.0 ? e[i] / ne : e[i]);
    }
    free(v); free(v_old); free(w); free(basis); free(alpha); free(beta);
    free(T); free(ev); free(V); free(e);
    return n_sel;
}

/* graph_to_negative_normalised_adjacency, src/data/init.rs:210-256: degrees
 * summed in COO order, M = -D^-1/2 A D^-1/2 in CSR (columns ascending), the
 * diagonal left out */
static void neg_norm_adj(int64_t n, const int32_t *r, const int32_t *c, const float *v,
                         int64_t nnz, int64_t **ip_out, int32_t **ix_out, double **dv_out) {
    double *deg = (double *)calloc((size_t)n, sizeof(double));
    for (int64_t e = 0; e < nnz; e++) deg[r[e]] += (double)v[e];
    double *dis = (double *)malloc(sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; i++) dis[i] = deg[i] > 1e-8 ? 1.0 / sqrt(deg[i]) : 0.0;
    /* CSR sorted by (row, col) */
    int64_t *ip = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t e = 0; e < nnz; e++) ip[r[e] + 1]++;
    for (int64_t i = 0; i < n; i++) ip[i + 1] += ip[i];
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    memcpy(cur, ip, sizeof(int64_t) * (size_t)n);
    int32_t *cc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    double *vv = (double *)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    for (int64_t e = 0; e < nnz; e++) {
        int64_t q = cur[r[e]]++;
        cc[q] = c[e];
        vv[q] = (double)v[e];
    }
    for (int64_t i = 0; i < n; i++) /* insertion sort of each row by column */
        for (int64_t a = ip[i] + 1; a < ip[i + 1]; a++)
            for (int64_t b = a; b > ip[i] && cc[b - 1] > cc[b]; b--) {
                int32_t tc = cc[b]; cc[b] = cc[b - 1]; cc[b - 1] = tc;
                double tv = vv[b]; vv[b] = vv[b - 1]; vv[b - 1] = tv;
            }
    int64_t *op = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    int32_t *oi = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    double *od = (double *)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    int64_t w = 0;
    for (int64_t i = 0; i < n; i++) {
        for (int64_t q = ip[i]; q < ip[i + 1]; q++)
            if (cc[q] != i) {
                oi[w] = cc[q];
                od[w] = -dis[i] * vv[q] * dis[cc[q]];
                w++;
            }
        op[i + 1] = w;
    }
    free(deg); free(dis); free(ip); free(cur); free(cc); free(vv);
    *ip_out = op;
    *ix_out = oi;
    *dv_out = od;
}

/* single_component_spectral_raw, src/data/init.rs:277-317: eigenvectors
 * 1..n_comp of M (random in [-1, 1) for a graph of <= n_comp + 1 vertices, or
 * for a dimension Lanczos did not deliver) into raw [n][n_comp] */
/* The sign of an eigenvector is free (the reference's follows its StdRng start
 * vector).  With a target `align` [n][n_comp] (test use: another
 * implementation's layout of the same vertices), each column of raw is
 * flipped to agree with it: centred dot product >= 0. */
static void align_signs(float *raw, const float *align, int64_t n, int n_comp) {
    if (!align) return;
    for (int d = 0; d < n_comp; d++) {
        double mr = 0.0, ma = 0.0, dot = 0.0;
        for (int64_t i = 0; i < n; i++) {
            mr += raw[i * n_comp + d];
            ma += align[i * n_comp + d];
        }
        mr /= (double)n;
        ma /= (double)n;
        for (int64_t i = 0; i < n; i++)
            dot += (raw[i * n_comp + d] - mr) * (align[i * n_comp + d] - ma);
        if (dot < 0.0)
            for (int64_t i = 0; i < n; i++) raw[i * n_comp + d] = -raw[i * n_comp + d];
    }
}

static void spectral_raw(int64_t n, const int32_t *r, const int32_t *c, const float *v,
                         int64_t nnz, int n_comp, uint64_t seed, float *raw) {
    if (n <= n_comp + 1) {
        for (int64_t i = 0; i < n * n_comp; i++) raw[i] = (float)(2.0 * u01(seed, 2, (uint64_t)i) - 1.0);
        return;
    }
    int64_t *ip;
    int32_t *ix;
    double *dv;
    neg_norm_adj(n, r, c, v, nnz, &ip, &ix, &dv);
    const int n_eigs = n_comp + 1 < n ? n_comp + 1 : (int)n;
    float *ev = (float *)malloc(sizeof(float) * (size_t)n * n_eigs);
    const int got = lanczos_smallest(n, ip, ix, dv, n_eigs, seed, ev);
    for (int d = 0; d < n_comp; d++)
        for (int64_t i = 0; i < n; i++)
            raw[i * n_comp + d] = d + 1 < got ? ev[i * n_eigs + d + 1]
                                              : (float)(2.0 * u01(seed + (uint64_t)d, 3, (uint64_t)i) - 1.0);
    free(ev); free(ip); free(ix); free(dv);
}

static float max_abs_f(const float *x, int64_t m) {
    float mx = 0.0f;
    for (int64_t i = 0; i < m; i++) {
        float a = fabsf(x[i]);
        if (a > mx) mx = a;
    }
    return mx;
}

/* centre each column (f32 sums in row order), scale max |x| to range */
static void centre_scale(float *emb, int64_t n, int n_comp, float range) {
    const float nt = (float)n;
    for (int d = 0; d < n_comp; d++) {
        float s = 0.0f;
        for (int64_t i = 0; i < n; i++) s += emb[i * n_comp + d];
        const float mean = s / nt;
        for (int64_t i = 0; i < n; i++) emb[i * n_comp + d] -= mean;
    }
    const float mx = max_abs_f(emb, n * n_comp);
    if (mx > 1e-8f) {
        const float sc = range / mx;
        for (int64_t i = 0; i < n * n_comp; i++) emb[i] *= sc;
    }
}

/* spectral_layout, src/data/init.rs:986-1004, with find_connected_components
 * (:330-391, BFS in ascending start vertex over the COO's rows),
 * single_component_spectral + finalise_spectral_embedding (:877-957) and
 * multi_component_init / component_meta_layout / meta_data_range
 * (:413-785).  graph: a symmetric COO of n vertices; data [n][d] optional
 * (NULL: simplex meta layout).  out [n][n_comp].  The reference's StdRng
 * draws (start vector, noise, random fills) are replaced by counter-based
 * draws of the same distributions. */
int orc_spectral_layout(int64_t n, const int32_t *rows, const int32_t *cols, const float *vals,
                        int64_t nnz, int32_t n_comp, uint64_t seed, float range, const float *data,
                        int32_t d, float *out) {
    return orc_spectral_layout_aligned(n, rows, cols, vals, nnz, n_comp, seed, range, data, d,
                                       NULL, out);
}

int orc_spectral_layout_aligned(int64_t n, const int32_t *rows, const int32_t *cols,
                                const float *vals, int64_t nnz, int32_t n_comp, uint64_t seed,
                                float range, const float *data, int32_t d, const float *align,
                                float *out) {
    if (n <= 0) return 1;
    /* adjacency in COO order per row, then BFS */
    int64_t *ip = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t e = 0; e < nnz; e++) ip[rows[e] + 1]++;
    for (int64_t i = 0; i < n; i++) ip[i + 1] += ip[i];
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    memcpy(cur, ip, sizeof(int64_t) * (size_t)n);
    int32_t *adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    for (int64_t e = 0; e < nnz; e++) adj[cur[rows[e]]++] = cols[e];
    int64_t *label = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *local = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)n); /* members, component by component */
    int64_t *cstart = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    for (int64_t i = 0; i < n; i++) label[i] = -1;
    int64_t nc = 0, filled = 0;
    for (int64_t s = 0; s < n; s++) {
        if (label[s] >= 0) continue;
        cstart[nc] = filled;
        int64_t head = filled;
        order[filled++] = s;
        label[s] = nc;
        while (head < filled) {
            int64_t node = order[head];
            local[node] = head - cstart[nc];
            head++;
            for (int64_t q = ip[node]; q < ip[node + 1]; q++)
                if (label[adj[q]] < 0) {
                    label[adj[q]] = nc;
                    order[filled++] = adj[q];
                }
        }
        nc++;
    }
    cstart[nc] = filled;
    if (nc == 1) {
        spectral_raw(n, rows, cols, vals, nnz, n_comp, seed, out);
        if (n > n_comp + 1) align_signs(out, align, n, n_comp);
        centre_scale(out, n, n_comp, range);
        for (int64_t i = 0; i < n * n_comp; i++)
            out[i] += (float)std_normal(seed + 9999, 4, (uint64_t)i) * 1e-4f;
    } else {
        /* component_meta_layout */
        float *meta = (float *)calloc((size_t)nc * n_comp, sizeof(float));
        int spectral_meta = nc > 2 * n_comp && data && nc <= 2000;
        if (nc > 2 * n_comp && data && nc > 2000) {
            for (int64_t l = 0; l < nc * n_comp; l++) meta[l] = (float)(2.0 * u01(seed, 5, (uint64_t)l) - 1.0);
        } else if (spectral_meta) {
            float *cen = (float *)calloc((size_t)nc * d, sizeof(float));
            for (int64_t l = 0; l < nc; l++) {
                const float inv = 1.0f / (float)(cstart[l + 1] - cstart[l]);
                for (int64_t q = cstart[l]; q < cstart[l + 1]; q++)
                    for (int f = 0; f < d; f++) cen[l * d + f] += data[order[q] * d + f];
                for (int f = 0; f < d; f++) cen[l * d + f] *= inv;
            }
            float *d2 = (float *)calloc((size_t)nc * nc, sizeof(float));
            float maxd2 = 0.0f;
            for (int64_t a = 0; a < nc; a++)
                for (int64_t b = a + 1; b < nc; b++) {
                    float s = 0.0f;
                    for (int f = 0; f < d; f++) {
                        float df = cen[a * d + f] - cen[b * d + f];
                        s += df * df;
                    }
                    d2[a * nc + b] = d2[b * nc + a] = s;
                    if (s > maxd2) maxd2 = s;
                }
            const float scale = maxd2 > 1e-12f ? 1.0f / maxd2 : 1.0f;
            int64_t m = nc * (nc - 1);
            int32_t *ar = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
            int32_t *ac = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
            float *av = (float *)malloc(sizeof(float) * (size_t)m);
            int64_t q = 0;
            for (int64_t a = 0; a < nc; a++)
                for (int64_t b = 0; b < nc; b++) {
                    if (a == b) continue;
                    ar[q] = (int32_t)a;
                    ac[q] = (int32_t)b;
                    av[q] = (float)exp(-(double)(d2[a * nc + b] * scale));
                    q++;
                }
            spectral_raw(nc, ar, ac, av, m, n_comp, seed, meta);
            if (align && nc > n_comp + 1) { /* target: the aligned layout's component centres */
                float *tc = (float *)calloc((size_t)nc * n_comp, sizeof(float));
                for (int64_t l = 0; l < nc; l++) {
                    for (int64_t q2 = cstart[l]; q2 < cstart[l + 1]; q2++)
                        for (int dd = 0; dd < n_comp; dd++) tc[l * n_comp + dd] += align[order[q2] * n_comp + dd];
                    for (int dd = 0; dd < n_comp; dd++) tc[l * n_comp + dd] /= (float)(cstart[l + 1] - cstart[l]);
                }
                align_signs(meta, tc, nc, n_comp);
                free(tc);
            }
            free(cen); free(d2); free(ar); free(ac); free(av);
        }
        float mx = max_abs_f(meta, nc * n_comp);
        if (!(nc > 2 * n_comp && data) || mx <= 1e-8f) { /* simplex */
            const int64_t k = (nc + 1) / 2;
            memset(meta, 0, sizeof(float) * (size_t)nc * n_comp);
            for (int64_t l = 0; l < nc; l++) {
                const int64_t idx = l < k ? l : l - k;
                if (idx < n_comp) meta[l * n_comp + idx] = l < k ? 1.0f : -1.0f;
            }
            mx = max_abs_f(meta, nc * n_comp);
        }
        if (mx > 1e-8f)
            for (int64_t l = 0; l < nc * n_comp; l++) meta[l] /= mx;
        /* each component around its centre */
        uint64_t draw = 0;
        for (int64_t l = 0; l < nc; l++) {
            const int64_t cn = cstart[l + 1] - cstart[l];
            float range_l = 0.0f; /* meta_data_range: half the largest distance to another centre */
            for (int64_t o = 0; o < nc; o++) {
                float s = 0.0f;
                for (int dd = 0; dd < n_comp; dd++) {
                    float df = meta[l * n_comp + dd] - meta[o * n_comp + dd];
                    s += df * df;
                }
                if (sqrtf(s) > range_l) range_l = sqrtf(s);
            }
            range_l /= 2.0f;
            if (cn < 2 * n_comp) {
                for (int64_t q = cstart[l]; q < cstart[l + 1]; q++)
                    for (int dd = 0; dd < n_comp; dd++)
                        out[order[q] * n_comp + dd] =
                            meta[l * n_comp + dd] + (float)(2.0 * u01(seed, 6, draw++) - 1.0) * range_l;
                continue;
            }
            /* the component's subgraph in COO order, vertices renamed to local */
            int64_t sm = 0;
            for (int64_t e = 0; e < nnz; e++) sm += label[rows[e]] == l && label[cols[e]] == l;
            int32_t *sr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(sm ? sm : 1));
            int32_t *sc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(sm ? sm : 1));
            float *sv = (float *)malloc(sizeof(float) * (size_t)(sm ? sm : 1));
            int64_t q = 0;
            for (int64_t e = 0; e < nnz; e++)
                if (label[rows[e]] == l && label[cols[e]] == l) {
                    sr[q] = (int32_t)local[rows[e]];
                    sc[q] = (int32_t)local[cols[e]];
                    sv[q] = vals[e];
                    q++;
                }
            float *sub = (float *)malloc(sizeof(float) * (size_t)cn * n_comp);
            spectral_raw(cn, sr, sc, sv, sm, n_comp, seed + (uint64_t)l, sub);
            if (align && cn > n_comp + 1) {
                float *ta = (float *)malloc(sizeof(float) * (size_t)cn * n_comp);
                for (int64_t lq = 0; lq < cn; lq++)
                    for (int dd = 0; dd < n_comp; dd++)
                        ta[lq * n_comp + dd] = align[order[cstart[l] + lq] * n_comp + dd];
                align_signs(sub, ta, cn, n_comp);
                free(ta);
            }
            const float smx = max_abs_f(sub, cn * n_comp);
            const float expansion = smx > 1e-8f ? range_l / smx : 1.0f;
            for (int64_t lq = 0; lq < cn; lq++)
                for (int dd = 0; dd < n_comp; dd++)
                    out[order[cstart[l] + lq] * n_comp + dd] =
                        meta[l * n_comp + dd] + sub[lq * n_comp + dd] * expansion;
            free(sr); free(sc); free(sv); free(sub);
        }
        centre_scale(out, n, n_comp, range);
        free(meta);
    }
    free(ip); free(cur); free(adj); free(label); free(local); free(order); free(cstart);
    return 0;
}

/* ------------------------------------------------------------------ PCA */

/* pca_layout, src/data/init.rs:1065-1140: centre the data (column means,
 * :1077-1085), the first n_comp principal-component scores U S (:1088-1114;
 * here X_c V with V the leading eigenvectors of the centred covariance, FP64,
 * by cyclic Jacobi), each score column centred and scaled to standard deviation
 * `range` (population variance; a column of std <= 1e-8 keeps scale 1,
 * :1117-1137).  Eigenvector signs: the entry of largest magnitude (the first
 * of equals) positive.  out [n][n_comp]. */
int orc_pca_layout(const float *x, int64_t n, int32_t d, int32_t n_comp, float range, float *out) {
    if (n <= 0) return 1;
    if (n_comp <= 0 || n_comp > d || n_comp > n) return 5;
    double *mean = (double *)calloc((size_t)d, sizeof(double));
    for (int64_t r = 0; r < n; r++)
        for (int j = 0; j < d; j++) mean[j] += (double)x[r * d + j];
    for (int j = 0; j < d; j++) mean[j] /= (double)n;
    double *cov = (double *)calloc((size_t)d * d, sizeof(double));
    double *row = (double *)malloc(sizeof(double) * (size_t)d);
    for (int64_t r = 0; r < n; r++) {
        for (int j = 0; j < d; j++) row[j] = (double)x[r * d + j] - mean[j];
        for (int i = 0; i < d; i++)
            for (int j = 0; j < d; j++) cov[i * d + j] += row[i] * row[j];
    }
    /* eigenvalues ascending, eigenvectors as columns (cyclic Jacobi above) */
    double *vec = (double *)malloc(sizeof(double) * (size_t)d * d);
    double *ev = (double *)malloc(sizeof(double) * (size_t)d);
    jacobi_eigen(cov, d, ev, vec);
    double *score = (double *)calloc((size_t)n, sizeof(double));
    for (int c = 0; c < n_comp; c++) {
        const int col = d - 1 - c; /* largest first */
        int big = 0;
        for (int i = 1; i < d; i++)
            if (fabs(vec[i * d + col]) > fabs(vec[big * d + col])) big = i;
        const double sg = vec[big * d + col] < 0.0 ? -1.0 : 1.0;
        double s = 0.0;
        for (int64_t r = 0; r < n; r++) {
            double t = 0.0;
            for (int j = 0; j < d; j++) t += ((double)x[r * d + j] - mean[j]) * sg * vec[j * d + col];
            score[r] = t;
            s += t;
        }
        const double mu = s / (double)n;
        double var = 0.0;
        for (int64_t r = 0; r < n; r++) var += (score[r] - mu) * (score[r] - mu);
        const double sd = sqrt(var / (double)n);
        const double sc = sd > 1e-8 ? (double)range / sd : 1.0;
        for (int64_t r = 0; r < n; r++) out[r * n_comp + c] = (float)((score[r] - mu) * sc);
    }
    free(mean); free(cov); free(row); free(vec); free(ev); free(score);
    return 0;
}
