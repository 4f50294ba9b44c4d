// The fit facades umap_gpu / densmap_gpu (src/lib.rs:3465-3487, :3829-3851,
// shared body umap_gpu_inner :3510-3708) over the device pipeline:
//
//   exhaustive kNN (MFMA shortlist + FP32 re-rank)  ->  fuzzy graph on the
//   device  ->  [densMAP: DensState on the device]  ->  initial layout  ->
//   device-resident Adam epochs  ->  [n_dim][n] output.
//
// Nothing returns to the host between stages; the only host syncs are the
// scalar size read-backs of graph construction.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <string>

#include "graph.hpp"
#include "knn.hpp"
#include "pca.hpp"
#include "spectral.hpp"
#include "fit.hpp"
#include "umap_state.hpp"

namespace mrs {

namespace {

constexpr int kWg = 256;

// counter-based uniform(-range, range): random_layout (src/data/init.rs:1024-1043)
// draws from StdRng; the stream is not reproducible here, the distribution is.
__global__ __launch_bounds__(kWg) void k_random_layout(float* __restrict__ out, uint64_t total,
                                                       uint64_t seed, float range) {
    uint64_t i = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (i >= total) return;
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + i * 0xBF58476D1CE4E5B9ull + 0x94D049BB133111EBull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
    out[i] = (float)((2.0 * u - 1.0) * (double)range);
}

// [n][dim] -> [dim][n] (the transpose at the end of umap_gpu_inner, :3698-3705)
__global__ __launch_bounds__(kWg) void k_transpose_out(const float* __restrict__ embd, uint32_t n,
                                                       int dim, float* __restrict__ out) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < dim; d++) out[(size_t)d * n + i] = embd[(size_t)i * dim + d];
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

std::string lower(const char* s) {
    std::string r = s ? s : "";
    for (auto& c : r) c = (char)std::tolower((unsigned char)c);
    return r;
}

}  // namespace

thread_local FitTimings g_last_fit;

// Where the calling thread's fits copy the kNN lists they computed (device
// buffers of the caller, capacity in entries; null: nowhere).
struct KnnCopyOut {
    int32_t* idx = nullptr;
    float* dist = nullptr;
    int64_t capacity = 0;
};
thread_local KnnCopyOut g_knn_out;

void random_layout_dev(mrs_ctx* ctx, float* out, int64_t n, int32_t dim, uint64_t seed,
                       float range) {
    const uint64_t total = (uint64_t)n * dim;
    hipLaunchKernelGGL(k_random_layout, dim3(div_up(total, kWg)), dim3(kWg), 0, ctx->stream, out,
                       total, seed, range);
    MRS_CHECK_LAUNCH();
}

std::unique_ptr<mrs_umap_state> fit_prepare(mrs_ctx* ctx, const float* data, int64_t n, int32_t d,
                                            const int32_t* pre_idx, const float* pre_dist,
                                            int32_t pre_k, const mrs_umap_params_gpu* p,
                                            const mrs_dens_params* dens, uint64_t seed,
                                            int32_t verbose, FitTimings& t) {
    if (!p) throw Error(MRS_INVALID_ARGUMENT, "params is NULL");
    if (n <= 0) throw Error(MRS_NO_DATA, "UMAP: empty data (NoData)");
    if (p->n_dim <= 0 || p->n_dim > 16) throw Error(MRS_INVALID_ARGUMENT, "n_dim must be 1..16");
    hipStream_t s = ctx->stream;

    // parse_umap_optimiser_gpu, src/lib.rs:3541-3547: unknown -> AdamGpu
    const std::string opt = lower(p->optimiser);
    if (opt == "adam_parallel" || opt == "adam" || opt == "sgd")
        throw Error(MRS_INVALID_ARGUMENT,
                    "optimiser '" + opt + "' is a CPU optimiser; this library runs \"adam_gpu\"");
    if (verbose > 0 && !opt.empty() && opt != "adam_gpu")
        std::printf("Unknown optimiser string provided (\"%s\"). Defaulting to GPU-accelerated Adam\n",
                    opt.c_str());
    if (p->dist_metric && lower(p->dist_metric) != "euclidean")
        throw Error(MRS_INVALID_ARGUMENT, "only the euclidean metric is implemented");

    // ---- kNN (run_ann_search_gpu) or precomputed
    DevView<float> xv(data, data ? (size_t)n * d : 0, s);
    DevBuf<int32_t> knn_idx_own;
    DevBuf<float> knn_dist_own;
    const int32_t* kidx;
    const float* kdist;
    int32_t k;
    std::unique_ptr<DevView<int32_t>> pre_iv;
    std::unique_ptr<DevView<float>> pre_dv;
    double t0 = now_ms();
    if (pre_idx && pre_dist) {
        if (pre_k <= 0) throw Error(MRS_INVALID_ARGUMENT, "precomputed k must be positive");
        k = pre_k;
        pre_iv.reset(new DevView<int32_t>(pre_idx, (size_t)n * k, s));
        pre_dv.reset(new DevView<float>(pre_dist, (size_t)n * k, s));
        kidx = pre_iv->p;
        kdist = pre_dv->p;
        if (verbose > 0) std::printf("Using precomputed kNN graph...\n");
    } else {
        if (!data) throw Error(MRS_NO_DATA, "UMAP: no data and no precomputed kNN (NoData)");
        k = p->k;
        const std::string ann = lower(p->ann_type);
        if (verbose > 0) {
            if (ann != "exhaustive_gpu" && ann != "exhaustive")
                std::printf("ann_type \"%s\": running the exact exhaustive_gpu search\n", ann.c_str());
            std::printf("Running GPU nearest neighbour search using exhaustive_gpu...\n");
        }
        knn_idx_own.alloc((size_t)n * k);
        knn_dist_own.alloc((size_t)n * k);
        KnnTimings kt;
        knn_exhaustive_dev(ctx, xv.p, n, d, k, p->knn_precision, p->knn_shortlist,
                           knn_idx_own.p, knn_dist_own.p, &kt);
        kidx = knn_idx_own.p;
        kdist = knn_dist_own.p;
        t.knn_shortlist_ms = kt.shortlist_ms;
        const int64_t entries = n * (int64_t)k;
        if (g_knn_out.idx && entries <= g_knn_out.capacity) {
            MRS_HIP(hipMemcpyAsync(g_knn_out.idx, kidx, sizeof(int32_t) * entries,
                                   hipMemcpyDeviceToDevice, s));
            MRS_HIP(hipMemcpyAsync(g_knn_out.dist, kdist, sizeof(float) * entries,
                                   hipMemcpyDeviceToDevice, s));
        }
    }
    MRS_HIP(hipStreamSynchronize(s));
    t.knn_ms = now_ms() - t0;

    // ---- graph
    t0 = now_ms();
    const mrs_umap_graph_params& gp = p->umap_graph_params;
    std::unique_ptr<mrs_graph> g(build_graph(ctx, kidx, kdist, n, k, gp.bandwidth,
                                             gp.local_connectivity, gp.mix_weight,
                                             p->optim_params.n_epochs));
    MRS_HIP(hipStreamSynchronize(s));
    t.graph_ms = now_ms() - t0;
    if (verbose > 0)
        std::printf("Finalised graph generation: %lld edges.\n", (long long)g->n_edges);

    // ---- densMAP state (DensState::new, src/lib.rs:3572-3589)
    DevBuf<float> dens_r, dens_mu;
    mrs_dens_state ds{};
    if (dens) {
        t0 = now_ms();
        dens_r.alloc((size_t)n);
        dens_mu.alloc((size_t)n);
        ds.params = *dens;
        ds.mu_tot = dens_state_on_device(g.get(), kidx, kdist, true, dens_r.p, dens_mu.p);
        ds.r = dens_r.p;
        ds.mu_sum = dens_mu.p;
        t.dens_ms = now_ms() - t0;
    }
    // the kNN lists are no longer needed (src/lib.rs:3591-3593)
    knn_idx_own.release();
    knn_dist_own.release();

    // ---- initial layout (initialise_embedding, src/data/init.rs:1166-1189)
    t0 = now_ms();
    const int32_t dim = p->n_dim;
    DevBuf<float> embd0((size_t)n * dim);
    const std::string init = lower(p->initialisation);
    const bool has_range = !std::isnan(p->init_range);
    // parse_initilisation (src/data/init.rs:75-85): "spectral", "random",
    // "pca"; any other string is PCA (src/lib.rs:3526-3540)
    if (init == "random") {
        random_layout_dev(ctx, embd0.p, n, dim, seed, has_range ? p->init_range : 10.0f);
    } else if (init == "spectral") {
        spectral_layout_dev(ctx, g.get(), dim, seed, has_range ? p->init_range : 10.0f,
                            embd0.p, xv.p, xv.p ? d : 0);
    } else {
        if (init != "pca" && verbose > 0)
            std::printf("Unknown initialisation provided: \"%s\". Defaulting to PCA.\n",
                        init.c_str());
        pca_layout_dev(ctx, xv.p, n, d, dim, has_range ? p->init_range : 1.0f, embd0.p);
    }
    MRS_HIP(hipStreamSynchronize(s));
    t.init_ms = now_ms() - t0;

    // ---- the optimiser state (optimise_embedding_adam_gpu's upload, umap_optimiser_gpu.rs:1302-1325)
    if (g->n_edges == 0) throw Error(MRS_NO_GRAPH_EDGES, "UMAP: no edges to optimise (NoGraphEdges)");
    mrs_umap_state* raw_state = nullptr;
    {
        int rc = mrs_umap_state_from_graph(ctx, g.get(), embd0.p, dim, &p->optim_params,
                                           dens ? &ds : nullptr, &raw_state);
        if (rc != MRS_OK) throw Error(rc, mrs_last_error());
    }
    t.n_edges = raw_state->n_edges;
    return std::unique_ptr<mrs_umap_state>(raw_state);
}

void fit_output_async(const mrs_umap_state* st, float* out_dev) {
    hipLaunchKernelGGL(k_transpose_out, dim3(div_up(st->n, kWg)), dim3(kWg), 0, st->ctx->stream,
                       st->embd.p, (uint32_t)st->n, st->dim, out_dev);
    MRS_CHECK_LAUNCH();
}

void fit_output(const mrs_umap_state* st, float* out) {
    hipStream_t s = st->ctx->stream;
    DevView<float> ov(out, (size_t)st->n * st->dim, s, false);
    hipLaunchKernelGGL(k_transpose_out, dim3(div_up(st->n, kWg)), dim3(kWg), 0, s, st->embd.p,
                       (uint32_t)st->n, st->dim, ov.p);
    MRS_CHECK_LAUNCH();
    ov.download(s);
    MRS_HIP(hipStreamSynchronize(s));
}

void umap_fit(mrs_ctx* ctx, const float* data, int64_t n, int32_t d, const int32_t* pre_idx,
              const float* pre_dist, int32_t pre_k, const mrs_umap_params_gpu* p,
              const mrs_dens_params* dens, uint64_t seed, int32_t verbose, float* out) {
    FitTimings t;
    const double t_all = now_ms();
    std::unique_ptr<mrs_umap_state> st =
        fit_prepare(ctx, data, n, d, pre_idx, pre_dist, pre_k, p, dens, seed, verbose, t);
    // ---- optimise (optimise_embedding_adam_gpu, umap_optimiser_gpu.rs:1302-1369)
    const double t0 = now_ms();
    if (verbose > 0)
        std::printf("Optimising embedding via Adam (GPU-accelerated) (%d epochs) on %lld edges...\n",
                    p->optim_params.n_epochs, (long long)st->n_edges);
    run_epochs(st.get(), 0, p->optim_params.n_epochs, (uint32_t)seed, p->epoch_strategy);
    MRS_HIP(hipStreamSynchronize(ctx->stream));
    t.epochs_ms = now_ms() - t0;
    // ---- output [n_dim][n]
    fit_output(st.get(), out);
    t.total_ms = now_ms() - t_all;
    g_last_fit = t;
    if (verbose > 0) std::printf("UMAP (GPU) complete!\n");
}

}  // namespace mrs

extern "C" {

void mrs_umap_params_gpu_default(mrs_umap_params_gpu* p) {
    if (!p) return;
    // UmapParamsGpu::default, src/lib.rs:3244-3260, with the exhaustive kNN
    p->n_dim = 2;
    p->k = 15;
    p->optimiser = "adam_gpu";
    p->ann_type = "exhaustive_gpu";
    p->initialisation = "spectral";
    p->init_range = NAN;
    p->dist_metric = "euclidean";
    p->knn_precision = MRS_KNN_FP32;
    p->knn_shortlist = 0;
    p->umap_graph_params.bandwidth = 1e-5f;
    p->umap_graph_params.local_connectivity = 1.0f;
    p->umap_graph_params.mix_weight = 1.0f;
    // UmapOptimParams::default_2d, src/training/umap_optimisers.rs:56-69
    p->optim_params.a = 1.5f;
    p->optim_params.b = 0.9f;
    p->optim_params.lr = 1.0f;
    p->optim_params.gamma = 1.0f;
    p->optim_params.n_epochs = 500;
    p->optim_params.neg_sample_rate = 5;
    p->optim_params.min_dist = 0.1f;
    p->optim_params.beta1 = 0.5f;
    p->optim_params.beta2 = 0.9f;
    p->optim_params.eps = 1e-7f;
    p->randomised = 0;
    p->epoch_strategy = MRS_EPOCH_EDGE_ATOMIC;
}

void mrs_umap_params_gpu_default_2d(float min_dist, float spread, mrs_umap_params_gpu* p) {
    if (!p) return;
    // UmapParamsGpu::new_default_2d, src/lib.rs:3338-3346: the defaults with
    // the optimiser's a, b fitted to min_dist (None: 0.5) and spread (None: 1.0)
    mrs_umap_params_gpu_default(p);
    const float md = std::isnan(min_dist) ? 0.5f : min_dist;
    const float sp = std::isnan(spread) ? 1.0f : spread;
    mrs_umap_optim_params o{};
    if (mrs_umap_optim_params_from_min_dist_spread(md, sp, NAN, NAN, -1, -1, NAN, NAN, NAN, &o) ==
        MRS_OK)
        p->optim_params = o;
}

int mrs_umap_gpu(mrs_ctx* ctx, const float* data, int64_t n, int32_t d, const int32_t* knn_idx,
                 const float* knn_dist, int32_t knn_k, const mrs_umap_params_gpu* params,
                 uint64_t seed, int32_t verbose, float* out) {
    return mrs::guarded([&] {
        if (!ctx || !out) throw mrs::Error(MRS_INVALID_ARGUMENT, "NULL argument");
        MRS_HIP(hipSetDevice(ctx->device));
        mrs::StreamScope scope(ctx->stream, ctx->pool);
        mrs::umap_fit(ctx, data, n, d, knn_idx, knn_dist, knn_k, params, nullptr, seed, verbose,
                      out);
    });
}

int mrs_densmap_gpu(mrs_ctx* ctx, const float* data, int64_t n, int32_t d,
                    const int32_t* knn_idx, const float* knn_dist, int32_t knn_k,
                    const mrs_umap_params_gpu* params, const mrs_dens_params* dens,
                    uint64_t seed, int32_t verbose, float* out) {
    return mrs::guarded([&] {
        if (!ctx || !out || !dens) throw mrs::Error(MRS_INVALID_ARGUMENT, "NULL argument");
        MRS_HIP(hipSetDevice(ctx->device));
        mrs::StreamScope scope(ctx->stream, ctx->pool);
        mrs::umap_fit(ctx, data, n, d, knn_idx, knn_dist, knn_k, params, dens, seed, verbose,
                      out);
    });
}

int mrs_fit_knn_copy_to(int32_t* idx, float* dist, int64_t capacity) {
    return mrs::guarded([&] {
        if ((idx == nullptr) != (dist == nullptr) || capacity < 0)
            throw mrs::Error(MRS_INVALID_ARGUMENT, "idx and dist: both or neither");
        if (idx && (!mrs::is_device_ptr(idx) || !mrs::is_device_ptr(dist)))
            throw mrs::Error(MRS_INVALID_ARGUMENT, "the kNN copy takes device buffers");
        mrs::g_knn_out.idx = idx;
        mrs::g_knn_out.dist = dist;
        mrs::g_knn_out.capacity = idx ? capacity : 0;
    });
}

int mrs_last_fit_timings(double* out) {
    return mrs::guarded([&] {
        if (!out) throw mrs::Error(MRS_INVALID_ARGUMENT, "NULL argument");
        const mrs::FitTimings& t = mrs::g_last_fit;
        out[0] = t.knn_ms;
        out[1] = t.graph_ms;
        out[2] = t.dens_ms;
        out[3] = t.init_ms;
        out[4] = t.epochs_ms;
        out[5] = t.total_ms;
        out[6] = (double)t.n_edges;
        out[7] = t.knn_shortlist_ms;
    });
}

}  // extern "C"
