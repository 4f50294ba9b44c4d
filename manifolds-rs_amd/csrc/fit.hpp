// The fit pipeline of umap_gpu_inner (src/lib.rs:3510-3708), split so the fit
// server can run the preparation on its workers and batch the epochs.
#pragma once

#include <memory>

#include "spectral.hpp"
#include "umap_state.hpp"

namespace mrs {

// kNN (or the precomputed lists) -> fuzzy graph -> [densMAP state] -> initial
// layout -> device optimiser state, on ctx's stream.  Fills the stage timings.
std::unique_ptr<mrs_umap_state> fit_prepare(mrs_ctx* ctx, const float* data, int64_t n, int32_t d,
                                            const int32_t* pre_idx, const float* pre_dist,
                                            int32_t pre_k, const mrs_umap_params_gpu* p,
                                            const mrs_dens_params* dens, uint64_t seed,
                                            int32_t verbose, FitTimings& t);

// The state's embedding as the reference's [n_dim][n] output (host or device
// pointer); synchronises the state's stream.
void fit_output(const mrs_umap_state* st, float* out);

// The same transpose into device memory, queued on the state's stream only.
void fit_output_async(const mrs_umap_state* st, float* out_dev);

void umap_fit(mrs_ctx* ctx, const float* data, int64_t n, int32_t d, const int32_t* pre_idx,
              const float* pre_dist, int32_t pre_k, const mrs_umap_params_gpu* p,
              const mrs_dens_params* dens, uint64_t seed, int32_t verbose, float* out);

}  // namespace mrs
