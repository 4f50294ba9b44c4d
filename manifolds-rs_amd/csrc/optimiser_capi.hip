// C ABI of the "adam_gpu" optimiser: state upload, the per-epoch launchers of
// src/training/umap_optimiser_gpu.rs:1016-1272 and the driver of :1302-1369.
#include <hip/hip_runtime.h>

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
        case 15: return {st->edge_ij.p, st->edge_ij.bytes()};
        default: throw Error(MRS_INVALID_ARGUMENT, "unknown state buffer id");
    }
}

}  // namespace

extern "C" {

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
        BufRef b = state_buffer(st, which);
        MRS_HIP(hipMemcpyAsync(outp, b.p, b.bytes, hipMemcpyDefault, st->ctx->stream));
        MRS_HIP(hipStreamSynchronize(st->ctx->stream));
    });
}

int mrs_umap_state_write(mrs_umap_state* st, int32_t which, const void* in) {
    return mrs::guarded([&] {
        if (!st || !in) throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
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
        MRS_HIP(hipStreamSynchronize(st->ctx->stream));
    });
}

int mrs_launch_adam_clear(mrs_umap_state* st, int32_t epoch) {
    return mrs::guarded([&] {
        if (!st) throw Error(MRS_INVALID_ARGUMENT, "state is NULL");
        mrs::launch_adam_clear(st, epoch, nullptr);
        st->grad_clean = true;
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
        st->grad_clean = true;
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
