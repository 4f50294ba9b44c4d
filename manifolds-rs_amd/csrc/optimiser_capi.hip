// C ABI of the "adam_gpu" optimiser: state upload, the per-epoch launchers of
// src/training/umap_optimiser_gpu.rs:1016-1272 and the driver of :1302-1369.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <vector>

#include "umap_state.hpp"

namespace {

using mrs::Error;

// Copy an adjacency list that may live in device memory to the host.
struct HostAdj {
    std::vector<int64_t> off;
    std::vector<int32_t> cols;
    std::vector<float> vals;
};

HostAdj fetch_adj(int64_t n, const int64_t* off, const int32_t* cols, const float* vals) {
    HostAdj h;
    h.off.resize((size_t)n + 1);
    MRS_HIP(hipMemcpy(h.off.data(), off, sizeof(int64_t) * (n + 1), hipMemcpyDefault));
    int64_t nnz = h.off[n];
    if (nnz < 0) throw Error(MRS_INVALID_ARGUMENT, "negative adjacency size");
    h.cols.resize((size_t)nnz);
    h.vals.resize((size_t)nnz);
    if (nnz) {
        MRS_HIP(hipMemcpy(h.cols.data(), cols, sizeof(int32_t) * nnz, hipMemcpyDefault));
        MRS_HIP(hipMemcpy(h.vals.data(), vals, sizeof(float) * nnz, hipMemcpyDefault));
    }
    return h;
}

void check_params(const mrs_umap_optim_params* p) {
    if (!p) throw Error(MRS_INVALID_ARGUMENT, "params is NULL");
    if (p->n_epochs <= 0) throw Error(MRS_INVALID_ARGUMENT, "n_epochs must be positive");
    if (p->neg_sample_rate < 0) throw Error(MRS_INVALID_ARGUMENT, "neg_sample_rate < 0");
}

struct BufRef {
    void* p;
    size_t bytes;
};

BufRef state_buffer(mrs_umap_state* st, int which) {
    switch (which) {
        case 0: return {st->embd.p, st->embd.bytes()};
        case 1: return {st->m.p, st->m.bytes()};
        case 2: return {st->v.p, st->v.bytes()};
        case 3: return {st->grad.p, st->grad.bytes()};
        case 4: return {st->has_update.p, st->has_update.bytes()};
        case 5: return {st->off.p, st->off.bytes()};
        case 6: return {st->csr_edge_idx.p, st->csr_edge_idx.bytes()};
        case 7: return {st->csr_other.p, st->csr_other.bytes()};
        case 8: return {st->per.p, st->per.bytes()};
        case 9: return {st->next.p, st->next.bytes()};
        case 10: return {st->mu.p, st->mu.bytes()};
        case 11: return {st->dens_re.p, st->dens_re.bytes()};
        case 12: return {st->dens_phi.p, st->dens_phi.bytes()};
        case 13: return {st->dens_w.p, st->dens_w.bytes()};
        case 14: return {st->dens_r.p, st->dens_r.bytes()};
        default: throw Error(MRS_INVALID_ARGUMENT, "unknown state buffer id");
    }
}

// UmapOptimParams::fit_params (src/training/umap_optimisers.rs:139-209): a, b of
// 1 / (1 + a x^(2b)) fitted by plain gradient descent (step 1, n_iter
// iterations, a clamped to [0.001, 10] and b to [0.1, 2] after each step) to
// the target curve 1 for x < min_dist, exp(-(x - min_dist) / spread) beyond,
// sampled at 300 points on [0, 3 spread].  Every operation is the reference's,
// in f32 and in its order (this file is built without contraction), so the
// result is the reference's to the last bit wherever libm's f32 powf / expf /
// logf agree.
void fit_curve_ab(float min_dist, float spread, int n_iter, float* a_out, float* b_out) {
    constexpr int kPoints = 300;
    const float two = 2.0f, one = 1.0f;
    const float step = spread * 3.0f / (float)(kPoints - 1);
    float xs[kPoints], ys[kPoints], lx[kPoints];
    for (int i = 0; i < kPoints; i++) {
        xs[i] = step * (float)i;
        ys[i] = xs[i] < min_dist ? one : expf(-(xs[i] - min_dist) / spread);
        lx[i] = xs[i] > 0.0f ? logf(xs[i]) : 0.0f;  // x = 0 takes no part in the fit
    }
    const float n_pts = (float)kPoints;
    float a = one, b = one;
    for (int it = 0; it < n_iter; it++) {
        float ga = 0.0f, gb = 0.0f;
        const float two_b = two * b;
        for (int i = 1; i < kPoints; i++) {  // xs[0] = 0 is skipped (x <= 0)
            const float p2b = powf(xs[i], two_b);
            const float den = one + a * p2b;
            const float den2 = den * den;
            const float err = one / den - ys[i];
            ga += err * (-p2b / den2);
            gb += err * (-two * a * p2b * lx[i] / den2);
        }
        ga /= n_pts;
        gb /= n_pts;
        a -= 1.0f * ga;
        b -= 1.0f * gb;
        a = std::min(std::max(a, 0.001f), 10.0f);
        b = std::min(std::max(b, 0.1f), 2.0f);
    }
    *a_out = a;
    *b_out = b;
}

}  // namespace

extern "C" {

int mrs_umap_optim_params_from_min_dist_spread(float min_dist, float spread, float lr,
                                               float gamma, int32_t n_epochs,
                                               int32_t neg_sample_rate, float beta1, float beta2,
                                               float eps, mrs_umap_optim_params* out) {
    return mrs::guarded([&] {
        if (!out) throw Error(MRS_INVALID_ARGUMENT, "out is NULL");
        if (!(spread > 0.0f) || std::isnan(min_dist))
            throw Error(MRS_INVALID_ARGUMENT, "spread must be positive and min_dist a number");
        // None (NaN / negative) takes the reference's default (:99-106)
        mrs_umap_optim_params p{};
        fit_curve_ab(min_dist, spread, 300, &p.a, &p.b);
        p.lr = std::isnan(lr) ? 1.0f : lr;
        p.gamma = std::isnan(gamma) ? 1.0f : gamma;
        p.n_epochs = n_epochs < 0 ? 500 : n_epochs;
        p.neg_sample_rate = neg_sample_rate < 0 ? 5 : neg_sample_rate;
        p.min_dist = min_dist;
        p.beta1 = std::isnan(beta1) ? 0.5f : beta1;   // UMAP_BETA1, src/training/mod.rs:16
        p.beta2 = std::isnan(beta2) ? 0.9f : beta2;   // UMAP_BETA2, :18
        p.eps = std::isnan(eps) ? 1e-7f : eps;        // EPS, :24
        *out = p;
    });
}

int mrs_umap_state_create(mrs_ctx* ctx, const float* embd, int64_t n, int32_t n_dim,
                          const int64_t* adj_off, const int32_t* adj_cols,
                          const float* adj_vals, const mrs_umap_optim_params* params,
                          const mrs_dens_state* dens, mrs_umap_state** out) {
    return mrs::guarded([&] {
        if (!ctx || !out) throw Error(MRS_INVALID_ARGUMENT, "ctx/out is NULL");
        check_params(params);
        mrs::StreamScope scope(ctx->stream, ctx->pool);
        if (n <= 0 || !embd) throw Error(MRS_NO_DATA, "UMAP: empty embedding (NoData)");
        if (n_dim <= 0 || n_dim > 16) throw Error(MRS_INVALID_ARGUMENT, "n_dim must be 1..16");
        if (n > (int64_t)INT32_MAX) throw Error(MRS_INVALID_ARGUMENT, "n too large");
        MRS_HIP(hipSetDevice(ctx->device));
        HostAdj h = fetch_adj(n, adj_off, adj_cols, adj_vals);
        mrs::HostCsr c = mrs::host_csr_from_graph(n, h.off.data(), h.cols.data(), h.vals.data());
        *out = mrs::make_state(ctx, embd, n, n_dim, c, *params, dens);
    });
}

int mrs_umap_state_destroy(mrs_umap_state* st) {
    return mrs::guarded([&] {
        if (!st) return;
        (void)hipStreamSynchronize(st->ctx->stream);
        mrs::StreamScope scope(st->ctx->stream, st->ctx->pool);
        delete st;
    });
}

int mrs_umap_state_info(const mrs_umap_state* st, int64_t* n, int32_t* n_dim, int64_t* n_edges) {
    return mrs::guarded([&] {
        if (!st) throw Error(MRS_INVALID_ARGUMENT, "state is NULL");
        if (n) *n = st->n;
        if (n_dim) *n_dim = st->dim;
        if (n_edges) *n_edges = st->n_edges;
    });
}

int mrs_umap_state_read(mrs_umap_state* st, int32_t which, void* outp) {
    return mrs::guarded([&] {
        if (!st || !outp) throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        if (which == 15 || which == 16) {
            // (i, j) or (slot under i, slot under j) of every 16-byte edge record
            const size_t half = 2 * sizeof(uint32_t);
            MRS_HIP(hipMemcpy2DAsync(outp, half,
                                     reinterpret_cast<const char*>(st->edges.p) +
                                         (which == 16 ? half : 0),
                                     sizeof(uint4), half, (size_t)st->n_edges, hipMemcpyDefault,
                                     st->ctx->stream));
            MRS_HIP(hipStreamSynchronize(st->ctx->stream));
            return;
        }
        BufRef b = state_buffer(st, which);
        MRS_HIP(hipMemcpyAsync(outp, b.p, b.bytes, hipMemcpyDefault, st->ctx->stream));
        MRS_HIP(hipStreamSynchronize(st->ctx->stream));
    });
}

int mrs_umap_state_write(mrs_umap_state* st, int32_t which, const void* in) {
    return mrs::guarded([&] {
        if (!st || !in) throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        if (which == 15 || which == 16)
            throw Error(MRS_INVALID_ARGUMENT, "the edge records follow the graph: read-only");
        BufRef b = state_buffer(st, which);
        MRS_HIP(hipMemcpyAsync(b.p, in, b.bytes, hipMemcpyDefault, st->ctx->stream));
        MRS_HIP(hipStreamSynchronize(st->ctx->stream));
        if (which == 3 || which == 4) st->grad_clean = false;
        if (which >= 11 && which <= 13) {  // a density buffer the edge kernel reads packed
            mrs::StreamScope scope(st->ctx->stream, st->ctx->pool);
            mrs::dens_pack_refresh(st);
            MRS_HIP(hipStreamSynchronize(st->ctx->stream));
        }
    });
}

int mrs_launch_grad_accum(mrs_umap_state* st, int32_t epoch, uint32_t seed,
                          const float* dens_coeff) {
    return mrs::guarded([&] {
        if (!st) throw Error(MRS_INVALID_ARGUMENT, "state is NULL");
        bool dens = dens_coeff != nullptr;
        if (dens) {
            if (!st->dens_enabled)
                throw Error(MRS_INVALID_ARGUMENT, "density coefficient given to a plain state");
            MRS_HIP(hipMemcpyAsync(st->dens_scalars.p + 3, dens_coeff, sizeof(float),
                                   hipMemcpyDefault, st->ctx->stream));
        }
        mrs::launch_grad_accum_node(st, epoch, seed, dens);
        MRS_HIP(hipStreamSynchronize(st->ctx->stream));
    });
}

int mrs_launch_dens_update(mrs_umap_state* st, float* coeff_out) {
    return mrs::guarded([&] {
        if (!st) throw Error(MRS_INVALID_ARGUMENT, "state is NULL");
        mrs::launch_dens_update_dev(st, true);
        float sc[4];
        MRS_HIP(hipMemcpyAsync(sc, st->dens_scalars.p, sizeof(sc), hipMemcpyDeviceToHost,
                               st->ctx->stream));
        MRS_HIP(hipStreamSynchronize(st->ctx->stream));
        if (coeff_out) *coeff_out = sc[3];
    });
}

int mrs_launch_adam_update(mrs_umap_state* st, int32_t epoch) {
    return mrs::guarded([&] {
        if (!st) throw Error(MRS_INVALID_ARGUMENT, "state is NULL");
        mrs::launch_adam_elem(st, epoch);
        MRS_HIP(hipStreamSynchronize(st->ctx->stream));
    });
}

int mrs_launch_edge_schedule_update(mrs_umap_state* st, int32_t epoch) {
    return mrs::guarded([&] {
        if (!st) throw Error(MRS_INVALID_ARGUMENT, "state is NULL");
        mrs::launch_schedule(st, epoch);
        MRS_HIP(hipStreamSynchronize(st->ctx->stream));
    });
}

int mrs_launch_grad_norm(mrs_umap_state* st, float* norm_out) {
    return mrs::guarded([&] {
        if (!st || !norm_out) throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        std::vector<float> g((size_t)st->n * st->dim);
        MRS_HIP(hipMemcpyAsync(g.data(), st->grad.p, st->grad.bytes(), hipMemcpyDeviceToHost,
                               st->ctx->stream));
        MRS_HIP(hipStreamSynchronize(st->ctx->stream));
        // host sum in node order, as the reference folds its [n] partials
        float sum = 0.0f;
        for (int64_t i = 0; i < st->n; i++) {
            float acc = 0.0f;
            for (int d = 0; d < st->dim; d++) {
                float x = g[(size_t)i * st->dim + d];
                acc += x * x;
            }
            sum += acc;
        }
        *norm_out = sqrtf(sum);
    });
}

int mrs_run_epochs(mrs_umap_state* st, int32_t first, int32_t count, uint64_t seed,
                   int32_t strategy) {
    return mrs::guarded([&] {
        if (!st) throw Error(MRS_INVALID_ARGUMENT, "state is NULL");
        if (first < 0 || count < 0 || first + count > st->params.n_epochs)
            throw Error(MRS_INVALID_ARGUMENT, "epoch range outside [0, n_epochs)");
        mrs::run_epochs(st, first, count, (uint32_t)seed, strategy);
    });
}

int mrs_optimise_embedding_adam_gpu(mrs_ctx* ctx, float* embd, int64_t n, int32_t n_dim,
                                    const int64_t* adj_off, const int32_t* adj_cols,
                                    const float* adj_vals,
                                    const mrs_umap_optim_params* params,
                                    const mrs_dens_state* dens, uint64_t seed,
                                    int32_t verbose, int32_t strategy) {
    mrs_umap_state* st = nullptr;
    int rc = mrs_umap_state_create(ctx, embd, n, n_dim, adj_off, adj_cols, adj_vals, params,
                                   dens, &st);
    if (rc != MRS_OK) return rc;
    rc = mrs::guarded([&] {
        if (verbose > 0)
            std::printf("Running %d epochs with GPU-accelerated Adam optimisation.\n",
                        params->n_epochs);
        // seed fits in u32; upper bits of a u64 seed are unused (reference :1326-1327)
        const uint32_t seed32 = (uint32_t)seed;
        const int32_t N = params->n_epochs;
        if (verbose > 0) {
            for (int32_t e = 0; e < N; e++) {
                mrs::run_epochs(st, e, 1, seed32, MRS_EPOCH_NODE_EXACT);
                if ((e + 1) % 100 == 0 || e + 1 == N) {
                    float gn = 0.0f;
                    if (mrs_launch_grad_norm(st, &gn) != MRS_OK)
                        throw Error(MRS_DEVICE_ERROR, mrs_last_error());
                    std::printf(" Epoch %d / %d: grad norm %.2f\n", e + 1, N, gn);
                }
            }
        } else {
            mrs::run_epochs(st, 0, N, seed32, strategy);
        }
        MRS_HIP(hipMemcpyAsync(embd, st->embd.p, st->embd.bytes(), hipMemcpyDefault,
                               ctx->stream));
        MRS_HIP(hipStreamSynchronize(ctx->stream));
    });
    mrs_umap_state_destroy(st);
    return rc;
}

}  // extern "C"

extern "C" {

int mrs_launch_epoch_edges(mrs_umap_state* st, int32_t epoch, uint32_t seed,
                           const float* dens_coeff) {
    return mrs::guarded([&] {
        if (!st) throw Error(MRS_INVALID_ARGUMENT, "state is NULL");
        bool dens = dens_coeff != nullptr;
        if (dens) {
            if (!st->dens_enabled)
                throw Error(MRS_INVALID_ARGUMENT, "density coefficient given to a plain state");
            MRS_HIP(hipMemcpyAsync(st->dens_scalars.p + 3, dens_coeff, sizeof(float),
                                   hipMemcpyDefault, st->ctx->stream));
        }
        mrs::launch_epoch_edges(st, epoch, seed, dens);
        mrs::derive_update_flags(st);  // has_update as the reference's kernel leaves it
        MRS_HIP(hipStreamSynchronize(st->ctx->stream));
    });
}

int mrs_launch_adam_clear(mrs_umap_state* st, int32_t epoch) {
    return mrs::guarded([&] {
        if (!st) throw Error(MRS_INVALID_ARGUMENT, "state is NULL");
        // flagged = has_update; gradients a caller wrote carry no markers yet
        if (!st->grad_clean) mrs::marks_from_flags(st);
        mrs::launch_adam_clear(st, epoch, nullptr);
        st->grad_clean = true;
        MRS_HIP(hipMemsetAsync(st->has_update.p, 0, st->has_update.bytes(), st->ctx->stream));
        MRS_HIP(hipStreamSynchronize(st->ctx->stream));
    });
}

}  // extern "C"

extern "C" {

int mrs_umap_state_set_profiling(mrs_umap_state* st, int32_t enable) {
    return mrs::guarded([&] {
        if (!st) throw Error(MRS_INVALID_ARGUMENT, "state is NULL");
        if (enable && !st->profiling) {
            mrs::prof_collect(st);
            for (int q = 0; q < 8; q++) {
                st->prof_ms[q] = 0.0;
                st->prof_count[q] = 0;
            }
        }
        st->profiling = enable != 0;
    });
}

int mrs_umap_state_kernel_times(mrs_umap_state* st, double* ms_out, int64_t* count_out) {
    return mrs::guarded([&] {
        if (!st) throw Error(MRS_INVALID_ARGUMENT, "state is NULL");
        mrs::prof_collect(st);
        for (int q = 0; q < 8; q++) {
            if (ms_out) ms_out[q] = st->prof_ms[q];
            if (count_out) count_out[q] = st->prof_count[q];
        }
    });
}

}  // extern "C"

extern "C" {

int mrs_umap_state_reset(mrs_umap_state* st, const float* embd) {
    return mrs::guarded([&] {
        if (!st || !embd) throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        hipStream_t s = st->ctx->stream;
        MRS_HIP(hipMemcpyAsync(st->embd.p, embd, st->embd.bytes(), hipMemcpyDefault, s));
        MRS_HIP(hipMemsetAsync(st->m.p, 0, st->m.bytes(), s));
        MRS_HIP(hipMemsetAsync(st->v.p, 0, st->v.bytes(), s));
        MRS_HIP(hipMemsetAsync(st->grad.p, 0, st->grad.bytes(), s));
        MRS_HIP(hipMemsetAsync(st->has_update.p, 0, st->has_update.bytes(), s));
        MRS_HIP(hipMemcpyAsync(st->next.p, st->per.p, st->per.bytes(), hipMemcpyDeviceToDevice, s));
        st->grad_clean = false;
        MRS_HIP(hipStreamSynchronize(s));
    });
}

int mrs_schedule_active_counts(const float* per, int64_t n_edges, int32_t n_epochs,
                               int64_t* counts) {
    return mrs::guarded([&] {
        if (n_edges < 0 || n_epochs < 0 || (!per && n_edges) || (!counts && n_epochs))
            throw Error(MRS_INVALID_ARGUMENT, "bad arguments");
        std::vector<float> p((size_t)n_edges);
        if (n_edges)
            MRS_HIP(hipMemcpy(p.data(), per, sizeof(float) * n_edges, hipMemcpyDefault));
        std::vector<float> next(p);
        for (int32_t e = 0; e < n_epochs; e++) {
            const float ef = (float)e;
            int64_t c = 0;
            for (int64_t q = 0; q < n_edges; q++)
                if (next[q] <= ef) {
                    next[q] += p[q];
                    c++;
                }
            counts[e] = c;
        }
    });
}

}  // extern "C"
