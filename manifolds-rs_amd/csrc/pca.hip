// PCA initial layout on the device: pca_layout (src/data/init.rs:1065-1140).
//
// The reference centres the data, takes a thin (or randomised) SVD of the
// centred matrix, forms the scores U S for the first n_comp components and
// scales each score column to mean 0 and standard deviation `range` (default
// PCA_RANGE = 1, :23).  Here the same scores come from the covariance route:
// U S = X_c V, where V holds the leading eigenvectors of X_c^T X_c.
//   1. column means of X, FP64 sums over a fixed grid (k_col_partial);
//   2. X_c^T X_c in FP64 from 64 x 64 tiles of the upper triangle, each
//      workgroup over a fixed chunk of rows, chunks summed in order
//      (k_cov_tiles, k_cov_finish): results identical run to run;
//   3. the d x d symmetric eigenproblem on the host: Householder
//      tridiagonalisation, then the library's implicit QL solver;
//   4. scores X_c v_c for the n_comp leading eigenvectors, one wavefront per
//      row with the lanes across the columns (k_scores), then the mean and
//      standard deviation of each column in FP64 and the scaling (k_scale).
// Eigenvector signs are not fixed by the mathematics (the reference's follow
// its SVD routine); here each eigenvector's entry of largest magnitude is made
// positive.  The randomised SVD of the reference approximates the same
// leading components; this path computes them exactly either way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "mrs_internal.hpp"
#include "pca.hpp"
#include "spectral.hpp"

namespace mrs {

void tridiag_eig(const std::vector<double>& alpha, const std::vector<double>& beta, int m,
                 std::vector<double>& evals, std::vector<double>& vecs);

namespace {

constexpr int kWg = 256;
constexpr int kTile = 64;      // covariance tile edge
constexpr int kTileRows = 16;  // rows staged per step
constexpr int kMaxComp = 16;

// part[block][c] = sum over the block's rows of x[r][c]
__global__ __launch_bounds__(kWg) void k_col_partial(const float* __restrict__ x, uint32_t n, int d,
                                                     uint32_t rows_per_block,
                                                     double* __restrict__ part) {
    const uint32_t r0 = blockIdx.x * rows_per_block;
    const uint32_t r1 = min(n, r0 + rows_per_block);
    for (int c = threadIdx.x; c < d; c += kWg) {
        double acc = 0.0;
        for (uint32_t r = r0; r < r1; r++) acc += (double)x[(size_t)r * d + c];
        part[(size_t)blockIdx.x * d + c] = acc;
    }
}

// Tile (ti, tj) of the upper triangle of X_c^T X_c over one chunk of rows:
// a 16 x 16 thread grid, each thread 4 x 4 entries in FP64; the chunk's rows
// are staged 16 at a time, centred in FP64, in LDS.
__global__ __launch_bounds__(kWg) void k_cov_tiles(const float* __restrict__ x, uint32_t n, int d,
                                                   const double* __restrict__ mean,
                                                   const uint2* __restrict__ tiles,
                                                   uint32_t rows_per_chunk,
                                                   double* __restrict__ part) {
    __shared__ double a[kTileRows][kTile];
    __shared__ double b[kTileRows][kTile];
    const uint2 t = tiles[blockIdx.x];
    const int c0i = (int)t.x * kTile, c0j = (int)t.y * kTile;
    const uint32_t r0 = blockIdx.y * rows_per_chunk;
    const uint32_t r1 = min(n, r0 + rows_per_chunk);
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[p][q] = 0.0;
    for (uint32_t rs = r0; rs < r1; rs += kTileRows) {
        // stage: 16 rows x 64 columns of each operand, 4 values per thread
        for (int e = threadIdx.x; e < kTileRows * kTile; e += kWg) {
            const int rr = e / kTile, cc = e % kTile;
            const uint32_t r = rs + rr;
            const int ci = c0i + cc, cj = c0j + cc;
            a[rr][cc] = (r < r1 && ci < d) ? (double)x[(size_t)r * d + ci] - mean[ci] : 0.0;
            b[rr][cc] = (r < r1 && cj < d) ? (double)x[(size_t)r * d + cj] - mean[cj] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int rr = 0; rr < kTileRows; rr++) {
            double av[4], bv[4];
#pragma unroll
            for (int p = 0; p < 4; p++) av[p] = a[rr][ty * 4 + p];
#pragma unroll
            for (int q = 0; q < 4; q++) bv[q] = b[rr][tx * 4 + q];
#pragma unroll
            for (int p = 0; p < 4; p++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc[p][q] = fma(av[p], bv[q], acc[p][q]);
        }
        __syncthreads();
    }
    double* out = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * kTile * kTile;
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < 4; q++) out[(ty * 4 + p) * kTile + tx * 4 + q] = acc[p][q];
}

// C[i][j] = C[j][i] = sum over the row chunks, in chunk order
__global__ __launch_bounds__(kWg) void k_cov_finish(const double* __restrict__ part,
                                                    const uint2* __restrict__ tiles,
                                                    uint32_t n_tiles, uint32_t chunks, int d,
                                                    double* __restrict__ cov) {
    const uint64_t e = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (e >= (uint64_t)n_tiles * kTile * kTile) return;
    const uint32_t tile = (uint32_t)(e / (kTile * kTile));
    const int w = (int)(e % (kTile * kTile));
    const int i = (int)tiles[tile].x * kTile + w / kTile, j = (int)tiles[tile].y * kTile + w % kTile;
    if (i >= d || j >= d || j < i) return;
    double s = 0.0;
    for (uint32_t c = 0; c < chunks; c++) s += part[((size_t)c * n_tiles + tile) * kTile * kTile + w];
    cov[(size_t)i * d + j] = s;
    cov[(size_t)j * d + i] = s;
}

// score[r][c] = sum_j (x[r][j] - mean[j]) v[c][j] for c < nc <= NC, one
// wavefront per row
template <int NC>
__global__ __launch_bounds__(kWg) void k_scores(const float* __restrict__ x, uint32_t n, int d,
                                                int nc, const double* __restrict__ mean,
                                                const double* __restrict__ v,
                                                double* __restrict__ score) {
    const int lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (kWg / 64);
    for (uint32_t r = blockIdx.x * (kWg / 64) + (threadIdx.x >> 6); r < n; r += waves) {
        double acc[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) acc[c] = 0.0;
        for (int j = lane; j < d; j += 64) {
            const double xv = (double)x[(size_t)r * d + j] - mean[j];
#pragma unroll
            for (int c = 0; c < NC; c++)
                if (c < nc) acc[c] = fma(xv, v[(size_t)c * d + j], acc[c]);
        }
#pragma unroll
        for (int c = 0; c < NC; c++) {
            double s = acc[c];
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
            if (lane == 0 && c < nc) score[(size_t)r * nc + c] = s;
        }
    }
}

// part[block][c] = sum over the block's rows of (score[r][c] - centre[c])^p, p = 1 or 2
__global__ __launch_bounds__(kWg) void k_score_moments(const double* __restrict__ score,
                                                       uint32_t n, int nc,
                                                       const double* __restrict__ centre,
                                                       uint32_t rows_per_block, int power,
                                                       double* __restrict__ part) {
    __shared__ double sh[kWg];
    const uint32_t r0 = blockIdx.x * rows_per_block;
    const uint32_t r1 = min(n, r0 + rows_per_block);
    for (int c = 0; c < nc; c++) {
        const double mu = centre ? centre[c] : 0.0;
        double acc = 0.0;
        for (uint32_t r = r0 + threadIdx.x; r < r1; r += kWg) {
            const double t = score[(size_t)r * nc + c] - mu;
            acc += power == 2 ? t * t : t;
        }
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int s = kWg / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * nc + c] = sh[0];
        __syncthreads();
    }
}

// out[r][c] = (score[r][c] - mean[c]) * scale[c]
__global__ __launch_bounds__(kWg) void k_scale(const double* __restrict__ score, uint64_t total,
                                               int nc, const double* __restrict__ mean_scale,
                                               float* __restrict__ out) {
    const uint64_t e = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (e >= total) return;
    const int c = (int)(e % nc);
    out[e] = (float)((score[e] - mean_scale[c]) * mean_scale[nc + c]);
}

// Householder reduction of the symmetric matrix a [m][m] (row-major, FP64) to
// tridiagonal form T = Q^T A Q (Golub & Van Loan, Matrix Computations, 4th
// ed., algorithm 8.3.1): diagonal alpha, off-diagonal beta, and Q [m][m].
void householder_tridiag(std::vector<double> a, int m, std::vector<double>& alpha,
                         std::vector<double>& beta, std::vector<double>& q) {
    q.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; i++) q[(size_t)i * m + i] = 1.0;
    std::vector<double> v(m), p(m), w(m);
    for (int k = 0; k + 2 < m; k++) {
        const int len = m - k - 1;
        double nx = 0.0;
        for (int i = 0; i < len; i++) nx += a[(size_t)(k + 1 + i) * m + k] * a[(size_t)(k + 1 + i) * m + k];
        nx = std::sqrt(nx);
        if (nx == 0.0) continue;
        const double x0 = a[(size_t)(k + 1) * m + k];
        const double al = x0 >= 0.0 ? -nx : nx;
        for (int i = 0; i < len; i++) v[i] = a[(size_t)(k + 1 + i) * m + k];
        v[0] -= al;
        double nv = 0.0;
        for (int i = 0; i < len; i++) nv += v[i] * v[i];
        nv = std::sqrt(nv);
        if (nv == 0.0) continue;
        for (int i = 0; i < len; i++) v[i] /= nv;
        // A22 <- H A22 H with H = I - 2 v v^T:  A22 -= v w^T + w v^T,
        // p = A22 v, K = v^T p, w = 2 (p - K v)
        double K = 0.0;
        for (int i = 0; i < len; i++) {
            double s = 0.0;
            const double* row = &a[(size_t)(k + 1 + i) * m + k + 1];
            for (int j = 0; j < len; j++) s += row[j] * v[j];
            p[i] = s;
            K += v[i] * s;
        }
        for (int i = 0; i < len; i++) w[i] = 2.0 * (p[i] - K * v[i]);
        for (int i = 0; i < len; i++) {
            double* row = &a[(size_t)(k + 1 + i) * m + k + 1];
            for (int j = 0; j < len; j++) row[j] -= v[i] * w[j] + w[i] * v[j];
        }
        a[(size_t)(k + 1) * m + k] = a[(size_t)k * m + k + 1] = al;
        for (int i = 1; i < len; i++) a[(size_t)(k + 1 + i) * m + k] = a[(size_t)k * m + k + 1 + i] = 0.0;
        // Q[:, k+1:] <- Q[:, k+1:] H
        for (int r = 0; r < m; r++) {
            double* row = &q[(size_t)r * m + k + 1];
            double s = 0.0;
            for (int j = 0; j < len; j++) s += row[j] * v[j];
            s *= 2.0;
            for (int j = 0; j < len; j++) row[j] -= s * v[j];
        }
    }
    alpha.resize(m);
    beta.assign(m > 1 ? m - 1 : 1, 0.0);
    for (int i = 0; i < m; i++) alpha[i] = a[(size_t)i * m + i];
    for (int i = 0; i + 1 < m; i++) beta[i] = a[(size_t)(i + 1) * m + i];
}

}  // namespace

void leading_eigenvectors(const std::vector<double>& cov, int d, int nc, std::vector<double>& vecs,
                          std::vector<double>& evals) {
    std::vector<double> alpha, beta, q, ev, z;
    householder_tridiag(cov, d, alpha, beta, q);
    tridiag_eig(alpha, beta, d, ev, z);  // ascending; z[k*d + r] = component k of eigenvector r
    vecs.assign((size_t)nc * d, 0.0);
    evals.assign(nc, 0.0);
    for (int c = 0; c < nc; c++) {
        const int r = d - 1 - c;  // largest first
        evals[c] = ev[r];
        double* out = &vecs[(size_t)c * d];
        for (int i = 0; i < d; i++) {
            double s = 0.0;
            for (int k = 0; k < d; k++) s += q[(size_t)i * d + k] * z[(size_t)k * d + r];
            out[i] = s;
        }
        // sign: the entry of largest magnitude (the first of equals) positive
        int big = 0;
        for (int i = 1; i < d; i++)
            if (std::fabs(out[i]) > std::fabs(out[big])) big = i;
        if (out[big] < 0.0)
            for (int i = 0; i < d; i++) out[i] = -out[i];
    }
}

void pca_layout_dev(mrs_ctx* ctx, const float* x, int64_t n64, int32_t d, int32_t nc,
                    float target_std, float* out) {
    if (n64 <= 0) throw Error(MRS_NO_DATA, "PCA: empty data (NoData)");
    if (!x) throw Error(MRS_INVALID_ARGUMENT, "PCA initialisation needs the data matrix");
    if (d <= 0 || nc <= 0 || nc > kMaxComp || nc > d || nc > n64)
        throw Error(MRS_INVALID_ARGUMENT, "PCA: n_comp must be 1..min(16, d, n)");
    if (d > 2048) throw Error(MRS_INVALID_ARGUMENT, "PCA: at most 2048 features");
    if (n64 > (int64_t)UINT32_MAX) throw Error(MRS_INVALID_ARGUMENT, "PCA: too many rows");
    const uint32_t n = (uint32_t)n64;
    hipStream_t s = ctx->stream;

    // 1. column means
    const uint32_t mb = std::min<uint32_t>(1024, div_up(n, 256));
    const uint32_t mrows = div_up(n, mb);
    DevBuf<double> mpart((size_t)mb * d), mean((size_t)d);
    hipLaunchKernelGGL(k_col_partial, dim3(mb), dim3(kWg), 0, s, x, n, d, mrows, mpart.p);
    MRS_CHECK_LAUNCH();
    std::vector<double> hm((size_t)mb * d), hmean((size_t)d, 0.0);
    MRS_HIP(hipMemcpyAsync(hm.data(), mpart.p, sizeof(double) * hm.size(), hipMemcpyDeviceToHost, s));
    MRS_HIP(hipStreamSynchronize(s));
    for (uint32_t b = 0; b < mb; b++)
        for (int c = 0; c < d; c++) hmean[c] += hm[(size_t)b * d + c];
    for (int c = 0; c < d; c++) hmean[c] /= (double)n;
    MRS_HIP(hipMemcpyAsync(mean.p, hmean.data(), sizeof(double) * d, hipMemcpyHostToDevice, s));

    // 2. X_c^T X_c, upper-triangle tiles
    const int nb = (d + kTile - 1) / kTile;
    std::vector<uint2> tl;
    for (int i = 0; i < nb; i++)
        for (int j = i; j < nb; j++) tl.push_back(make_uint2((unsigned)i, (unsigned)j));
    const uint32_t nt = (uint32_t)tl.size();
    const uint32_t chunks = std::max<uint32_t>(
        1, std::min<uint32_t>(div_up(n, 2048), std::max<uint32_t>(1, 1024 / nt)));
    const uint32_t crows = div_up(n, chunks);
    DevBuf<uint2> dtiles(nt);
    DevBuf<double> cpart((size_t)chunks * nt * kTile * kTile), cov((size_t)d * d);
    MRS_HIP(hipMemcpyAsync(dtiles.p, tl.data(), sizeof(uint2) * nt, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_cov_tiles, dim3(nt, chunks), dim3(kWg), 0, s, x, n, d, mean.p, dtiles.p,
                       crows, cpart.p);
    MRS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_cov_finish, dim3(div_up((uint64_t)nt * kTile * kTile, kWg)), dim3(kWg), 0,
                       s, cpart.p, dtiles.p, nt, chunks, d, cov.p);
    MRS_CHECK_LAUNCH();
    std::vector<double> hcov((size_t)d * d);
    MRS_HIP(hipMemcpyAsync(hcov.data(), cov.p, sizeof(double) * hcov.size(), hipMemcpyDeviceToHost, s));
    MRS_HIP(hipStreamSynchronize(s));

    // 3. leading eigenvectors on the host
    std::vector<double> vecs, evals;
    leading_eigenvectors(hcov, d, nc, vecs, evals);
    DevBuf<double> dv(vecs.size());
    MRS_HIP(hipMemcpyAsync(dv.p, vecs.data(), sizeof(double) * vecs.size(), hipMemcpyHostToDevice, s));

    // 4. scores, their mean and standard deviation, the scaling
    DevBuf<double> score((size_t)n * nc);
    const unsigned sb = std::min<unsigned>(4096, div_up(n, kWg / 64));
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(sb), dim3(kWg), 0, s, x, n, d, nc, mean.p, dv.p, score.p);
    };
    if (nc <= 2) go(k_scores<2>);
    else if (nc <= 4) go(k_scores<4>);
    else if (nc <= 8) go(k_scores<8>);
    else go(k_scores<kMaxComp>);
    MRS_CHECK_LAUNCH();
    std::vector<double> hp((size_t)mb * nc);
    DevBuf<double> spart((size_t)mb * nc), ms((size_t)2 * nc);
    std::vector<double> hms((size_t)2 * nc, 0.0);
    hipLaunchKernelGGL(k_score_moments, dim3(mb), dim3(kWg), 0, s, score.p, n, nc, nullptr, mrows, 1,
                       spart.p);
    MRS_CHECK_LAUNCH();
    MRS_HIP(hipMemcpyAsync(hp.data(), spart.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
    MRS_HIP(hipStreamSynchronize(s));
    for (uint32_t b = 0; b < mb; b++)
        for (int c = 0; c < nc; c++) hms[c] += hp[(size_t)b * nc + c];
    for (int c = 0; c < nc; c++) hms[c] /= (double)n;
    MRS_HIP(hipMemcpyAsync(ms.p, hms.data(), sizeof(double) * nc, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_score_moments, dim3(mb), dim3(kWg), 0, s, score.p, n, nc, ms.p, mrows, 2,
                       spart.p);
    MRS_CHECK_LAUNCH();
    MRS_HIP(hipMemcpyAsync(hp.data(), spart.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
    MRS_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < nc; c++) {
        double var = 0.0;
        for (uint32_t b = 0; b < mb; b++) var += hp[(size_t)b * nc + c];
        const double sd = std::sqrt(var / (double)n);
        hms[nc + c] = sd > 1e-8 ? (double)target_std / sd : 1.0;
    }
    MRS_HIP(hipMemcpyAsync(ms.p, hms.data(), sizeof(double) * 2 * nc, hipMemcpyHostToDevice, s));
    const uint64_t total = (uint64_t)n * nc;
    hipLaunchKernelGGL(k_scale, dim3(div_up(total, kWg)), dim3(kWg), 0, s, score.p, total, nc, ms.p,
                       out);
    MRS_CHECK_LAUNCH();
    MRS_HIP(hipStreamSynchronize(s));
}

}  // namespace mrs

extern "C" int mrs_pca_layout(mrs_ctx* ctx, const float* data, int64_t n, int32_t d,
                              int32_t n_comp, int32_t randomised, float range, uint64_t seed,
                              float* out) {
    (void)randomised;
    (void)seed;
    return mrs::guarded([&] {
        if (!ctx || !out) throw mrs::Error(MRS_INVALID_ARGUMENT, "NULL argument");
        if (!data) throw mrs::Error(MRS_INVALID_ARGUMENT, "PCA initialisation needs the data matrix");
        if (n <= 0) throw mrs::Error(MRS_NO_DATA, "PCA: empty data (NoData)");
        if (d <= 0) throw mrs::Error(MRS_INVALID_ARGUMENT, "PCA: d must be positive");
        mrs::StreamScope scope(ctx->stream, ctx->pool);
        hipStream_t s = ctx->stream;
        mrs::DevView<float> xv(data, (size_t)n * d, s);
        mrs::DevView<float> ov(out, (size_t)n * (n_comp > 0 ? n_comp : 0), s, false);
        mrs::pca_layout_dev(ctx, xv.p, n, d, n_comp, std::isnan(range) ? 1.0f : range, ov.p);
        ov.download(s);
        MRS_HIP(hipStreamSynchronize(s));
    });
}

extern "C" int mrs_sym_eigen_leading(const double* a, int32_t d, int32_t n_comp, double* vecs,
                                     double* evals) {
    return mrs::guarded([&] {
        if (!a || !vecs || !evals || d <= 0 || n_comp <= 0 || n_comp > d)
            throw mrs::Error(MRS_INVALID_ARGUMENT, "bad symmetric eigenproblem");
        std::vector<double> m(a, a + (size_t)d * d), v, e;
        mrs::leading_eigenvectors(m, d, n_comp, v, e);
        std::copy(v.begin(), v.end(), vecs);
        std::copy(e.begin(), e.end(), evals);
    });
}

