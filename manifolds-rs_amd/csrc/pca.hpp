// PCA initial layout on the device (pca_layout, src/data/init.rs:1065-1140).
#pragma once

#include <cstdint>
#include <vector>

#include "mrs_internal.hpp"

namespace mrs {

// out [n][nc] (device) = the data's first nc principal-component scores, each
// column centred and scaled to standard deviation target_std.  x: device
// [n][d] f32.  Synchronises the context's stream.
void pca_layout_dev(mrs_ctx* ctx, const float* x, int64_t n, int32_t d, int32_t nc,
                    float target_std, float* out);

// The nc leading eigenvectors (largest eigenvalue first) of the symmetric
// matrix cov [d][d] as rows of vecs [nc][d], each with its entry of largest
// magnitude positive; evals [nc].  Host.
void leading_eigenvectors(const std::vector<double>& cov, int d, int nc, std::vector<double>& vecs,
                          std::vector<double>& evals);

}  // namespace mrs
