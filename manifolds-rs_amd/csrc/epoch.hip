// The "adam_gpu" layout epoch on CDNA4 (gfx950).
//
// Two strategies share one device state:
//
//  * MRS_EPOCH_EDGE_ATOMIC (default): one thread per undirected edge over a
//    coalesced (i, j, next, per) edge stream.  An active edge computes the
//    attractive (and densMAP) term once, applies it with opposite signs to both
//    endpoints, draws `neg_rate` hashed negatives for each endpoint exactly as
//    the reference's per-CSR-entry walk does (same (node, edge_local) keys), and
//    scatters the per-endpoint sums with FP32 atomics.  The smaller endpoint's
//    contributions are first summed across the wavefront (edges are sorted by
//    i, so runs of equal i are contiguous), cutting its atomics by the run
//    length.  The schedule advance is fused (the edge thread owns next[e]).
//    A node kernel then applies Adam to flagged nodes and clears grad/flags.
//
//  * MRS_EPOCH_NODE_EXACT: the reference's node-parallel CSR walk
//    (umap_grad_accum, src/training/umap_optimiser_gpu.rs:682-860), summing in
//    the reference's order; deterministic bit-for-bit run to run.  Followed by
//    the element-wise Adam kernel (:887-923) and the schedule kernel (:940-955).
//
// densMAP (both strategies): umap_dens_radii (:487-544) as a node kernel, then
// the correlation statistics of src/utils/density.rs:464-508 reduced in FP64 on
// the device with a fixed grid (deterministic), then umap_dens_weights
// (:568-586) fused with the final reduction.  No host readback per epoch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "autotune.hpp"
#include "epoch_device.hpp"
#include "umap_state.hpp"

namespace mrs {

namespace {

constexpr int kWg = 256;

// ---------------------------------------------------------------- helpers

template <int DIM>
struct Dim {
    static constexpr int value = DIM;
};

// ---------------------------------------------------- node-exact gradient

// umap_grad_accum, umap_optimiser_gpu.rs:682-860.  DIM == 0: run-time n_dim <= 16.
template <int DIM, bool DENS>
__global__ __launch_bounds__(kWg) void k_grad_accum_node(
    const float* __restrict__ embd, const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ eidx, const uint32_t* __restrict__ other_node,
    const float* __restrict__ next, float* __restrict__ grad, uint32_t* __restrict__ has_update,
    uint32_t n, int dim_rt, uint32_t epoch, float epoch_f, uint32_t seed, FastMod fm, float a,
    float b, float two_a_b, float two_gamma_b, int neg_rate, const float* __restrict__ mu,
    const float* __restrict__ dens_re, const float* __restrict__ dens_phi,
    const float* __restrict__ dens_w, const float* __restrict__ dens_scalars) {
    constexpr int MAXD = DIM > 0 ? DIM : 16;
    const int nd = DIM > 0 ? DIM : dim_rt;
    uint32_t node = blockIdx.x * kWg + threadIdx.x;
    if (node >= n) return;
    const float two = 2.0f, one = 1.0f;
    float ys[MAXD], g[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; d++) {
        ys[d] = d < nd ? embd[(size_t)node * nd + d] : 0.0f;
        g[d] = 0.0f;
    }
    float coeff = 0.0f, re_self = 0.0f, phi_self = 0.0f, w_self = 0.0f;
    if (DENS) {
        coeff = dens_scalars[3];
        re_self = dens_re[node];
        phi_self = dens_phi[node];
        w_self = dens_w[node];
    }
    uint32_t active = 0, edge_local = 0;
    const uint32_t start = off[node], end = off[node + 1];
    for (uint32_t pos = start; pos < end; pos++, edge_local++) {
        const uint32_t e = eidx[pos];
        if (next[e] > epoch_f) continue;
        active = 1;
        const uint32_t other = other_node[pos];
        float yo[MAXD];
        float dist_sq = 0.0f;
#pragma unroll
        for (int d = 0; d < MAXD; d++) {
            yo[d] = d < nd ? embd[(size_t)other * nd + d] : 0.0f;
            if (d < nd) {
                float diff = ys[d] - yo[d];
                dist_sq += diff * diff;
            }
        }
        if (dist_sq >= kGradDistSqThreshold) {
            float dsb = powf(dist_sq, b);
            float denom = one + a * dsb;
            float gc = two_a_b * dsb / (dist_sq * denom);
#pragma unroll
            for (int d = 0; d < MAXD; d++)
                if (d < nd) g[d] += two * gc * (yo[d] - ys[d]);
            if (DENS) {
                float phi = one / denom;
                float dphi_term = a * b * dsb / dist_sq * phi;
                float common = one - b * (one - phi);
                float dr_self = (phi / phi_self) * (common / expf(re_self) + dphi_term);
                float dr_other =
                    (phi / dens_phi[other]) * (common / expf(dens_re[other]) + dphi_term);
                float cor = coeff * (w_self * dr_self + dens_w[other] * dr_other) / mu[e];
                float two_cor = two * cor;
#pragma unroll
                for (int d = 0; d < MAXD; d++)
                    if (d < nd) g[d] += two * clampf(two_cor * (ys[d] - yo[d]), kGradClipVal);
            }
        }
        for (int s = 0; s < neg_rate; s++) {
            uint32_t k = hash_neg(seed, node, epoch, edge_local, (uint32_t)s, fm);
            if (k == node) continue;
            float yk[MAXD];
            float dsk = 0.0f;
#pragma unroll
            for (int d = 0; d < MAXD; d++) {
                yk[d] = d < nd ? embd[(size_t)k * nd + d] : 0.0f;
                if (d < nd) {
                    float diff = ys[d] - yk[d];
                    dsk += diff * diff;
                }
            }
            float safe = dsk + kGradRepEps;
            float dsb = powf(safe, b);
            float denom = safe * (one + a * dsb);
            float gc = clampf(two_gamma_b / denom, kGradClipVal);
#pragma unroll
            for (int d = 0; d < MAXD; d++)
                if (d < nd) g[d] += gc * (ys[d] - yk[d]);
        }
    }
#pragma unroll
    for (int d = 0; d < MAXD; d++)
        if (d < nd) grad[(size_t)node * nd + d] = g[d];
    has_update[node] = active;
}

// umap_adam_update, umap_optimiser_gpu.rs:887-923 (one thread per element)
__global__ __launch_bounds__(kWg) void k_adam_elem(const float* __restrict__ grad,
                                                   const uint32_t* __restrict__ has_update,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   float* __restrict__ embd, uint64_t n_total,
                                                   int dim, float fused_lr, float epsc,
                                                   float omb1, float omb2) {
    uint64_t i = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (i >= n_total) return;
    if (has_update[i / dim] == 0) return;
    float g = grad[i];
    float mo = m[i];
    float mn = mo + omb1 * (g - mo);
    m[i] = mn;
    float vo = v[i];
    float vn = vo + omb2 * (g * g - vo);
    v[i] = vn;
    embd[i] += fused_lr * mn / (sqrtf(vn) + epsc);
}

// umap_edge_schedule_update, umap_optimiser_gpu.rs:940-955
__global__ __launch_bounds__(kWg) void k_schedule(const float* __restrict__ per,
                                                  float* __restrict__ next, uint64_t n_edges,
                                                  float epoch_f) {
    uint64_t e = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (e >= n_edges) return;
    float nx = next[e];
    if (nx <= epoch_f) next[e] = nx + per[e];
}

// umap_grad_norm_sq, umap_optimiser_gpu.rs:976-997
__global__ __launch_bounds__(kWg) void k_grad_norm_sq(const float* __restrict__ grad,
                                                      float* __restrict__ partial, uint32_t n,
                                                      int dim) {
    uint32_t node = blockIdx.x * kWg + threadIdx.x;
    if (node >= n) return;
    float acc = 0.0f;
    for (int d = 0; d < dim; d++) {
        float g = grad[(size_t)node * dim + d];
        acc += g * g;
    }
    partial[node] = acc;
}

// ----------------------------------------------------------------- densMAP

// umap_dens_radii, umap_optimiser_gpu.rs:487-544
template <int DIM>
__global__ __launch_bounds__(kWg) void k_dens_radii(const float* __restrict__ embd,
                                                    const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ other_node,
                                                    float* __restrict__ re,
                                                    float* __restrict__ phi_sum, uint32_t n,
                                                    int dim_rt, float a, float b,
                                                    float dens_eps, float tiny) {
    constexpr int MAXD = DIM > 0 ? DIM : 16;
    const int nd = DIM > 0 ? DIM : dim_rt;
    uint32_t node = blockIdx.x * kWg + threadIdx.x;
    if (node >= n) return;
    float ys[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; d++) ys[d] = d < nd ? embd[(size_t)node * nd + d] : 0.0f;
    float sum_sq = 0.0f, sum_phi = 0.0f;
    for (uint32_t pos = off[node]; pos < off[node + 1]; pos++) {
        uint32_t o = other_node[pos];
        float dist_sq = 0.0f;
#pragma unroll
        for (int d = 0; d < MAXD; d++)
            if (d < nd) {
                float diff = ys[d] - embd[(size_t)o * nd + d];
                dist_sq += diff * diff;
            }
        float dsb = powf(dist_sq, b);
        float phi = 1.0f / (1.0f + a * dsb);
        sum_sq += phi * dist_sq;
        sum_phi += phi;
    }
    float denom = sum_phi < tiny ? tiny : sum_phi;
    re[node] = logf(dens_eps + sum_sq / denom);
    phi_sum[node] = sum_phi;
}

// The same radii with 16 lanes per node (lane l takes the row's entries l,
// l+16, ...; the partial sums meet in a fixed xor tree, so the result is
// identical run to run).  A node's ~2k entries are random gathers: one thread
// per node leaves them mostly waiting in sequence, 16 lanes overlap them.
// Used by the edge-parallel strategy; the node-exact one keeps the reference's
// sequential order above.
constexpr int kRadiiLanes = 16;
template <int DIM>
__global__ __launch_bounds__(kWg) void k_dens_radii_lanes(const float* __restrict__ embd,
                                                          const uint32_t* __restrict__ off,
                                                          const uint32_t* __restrict__ other_node,
                                                          const uint32_t* __restrict__ outmap,
                                                          float* __restrict__ re,
                                                          float* __restrict__ phi_sum, uint32_t n,
                                                          float a, float b, float dens_eps,
                                                          float tiny) {
    const uint32_t node = (blockIdx.x * kWg + threadIdx.x) / kRadiiLanes;
    const int sub = threadIdx.x & (kRadiiLanes - 1);
    float sum_sq = 0.0f, sum_phi = 0.0f;
    if (node < n) {
        float ys[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) ys[d] = embd[(size_t)node * DIM + d];
        auto add = [&](const float (&yo)[DIM]) {
            float dist_sq = 0.0f;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                float diff = ys[d] - yo[d];
                dist_sq += diff * diff;
            }
            // a neighbour at the node's own position (dist_sq = 0) has phi = 1
            float dsb = dist_sq > 0.0f ? pos_powf(dist_sq, b) : 0.0f;
            float phi = 1.0f / (1.0f + a * dsb);
            sum_sq += phi * dist_sq;
            sum_phi += phi;
        };
        // a lane's entries pos, pos + 16, ... are added in that order; four at
        // a time have their rows gathered together first, so four gathers are
        // in flight per lane instead of one after another
        constexpr int kU = 4;
        const uint32_t end = off[node + 1];
        uint32_t pos = off[node] + sub;
        for (; pos + (kU - 1) * kRadiiLanes < end; pos += kU * kRadiiLanes) {
            uint32_t o[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) o[u] = other_node[pos + u * kRadiiLanes];
            float yo[kU][DIM];
#pragma unroll
            for (int u = 0; u < kU; u++)
#pragma unroll
                for (int d = 0; d < DIM; d++) yo[u][d] = embd[(size_t)o[u] * DIM + d];
#pragma unroll
            for (int u = 0; u < kU; u++) add(yo[u]);
        }
        for (; pos < end; pos += kRadiiLanes) {
            const uint32_t o = other_node[pos];
            float yo[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) yo[d] = embd[(size_t)o * DIM + d];
            add(yo);
        }
    }
#pragma unroll
    for (int o = kRadiiLanes / 2; o >= 1; o >>= 1) {
        sum_sq += __shfl_xor(sum_sq, o, kRadiiLanes);
        sum_phi += __shfl_xor(sum_phi, o, kRadiiLanes);
    }
    if (node < n && sub == 0) {
        const uint32_t out = outmap ? outmap[node] : node;
        float denom = sum_phi < tiny ? tiny : sum_phi;
        re[out] = logf(dens_eps + sum_sq / denom);
        phi_sum[out] = sum_phi;
    }
}

// ---- densMAP locality order
// The radii kernel gathers every neighbour's embedding row.  In node order
// those rows are scattered over the whole embedding; in an order where graph
// neighbours sit close together most of them are L2 hits.  The order is a
// Z-order (Morton) sort of the first min(dim, 2) coordinates of the embedding
// at the state's first density update (graph neighbours lie close together in
// the layout by then), built once per state.  Each density epoch gathers the embedding into that
// order and runs the radii kernel on the CSR renumbered into it, writing each
// radius back to its node.  Every row keeps its entries in their order, so
// every node adds the same terms in the same order: the radii are identical
// to those of node order.
__device__ inline uint32_t ordered_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// bbox[2d] / bbox[2d+1]: ordered-bit minimum / maximum of coordinate d < nd
__global__ __launch_bounds__(kWg) void k_embd_bbox(const float* __restrict__ embd, uint32_t n,
                                                   int dim, int nd, uint32_t* __restrict__ bbox) {
    uint32_t lo[2] = {~0u, ~0u}, hi[2] = {0u, 0u};
    for (uint32_t i = blockIdx.x * kWg + threadIdx.x; i < n; i += gridDim.x * kWg)
        for (int d = 0; d < nd; d++) {
            const uint32_t o = ordered_bits(embd[(size_t)i * dim + d]);
            lo[d] = o < lo[d] ? o : lo[d];
            hi[d] = o > hi[d] ? o : hi[d];
        }
    for (int d = 0; d < nd; d++) {
        atomicMin(&bbox[2 * d], lo[d]);
        atomicMax(&bbox[2 * d + 1], hi[d]);
    }
}

__device__ inline uint32_t spread_bits16(uint32_t x) {  // bit k -> bit 2k
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// key[i] = Z-order of node i's first nd coordinates, each quantised to 16 bits
__global__ __launch_bounds__(kWg) void k_morton_keys(const float* __restrict__ embd, uint32_t n,
                                                     int dim, int nd, float lo0, float sc0,
                                                     float lo1, float sc1,
                                                     uint32_t* __restrict__ key) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    auto quant = [](float x, float lo, float sc) {
        const float t = (x - lo) * sc;
        return t > 0.0f ? (t < 65535.0f ? (uint32_t)t : 65535u) : 0u;  // NaN -> 0
    };
    const uint32_t q0 = quant(embd[(size_t)i * dim], lo0, sc0);
    key[i] = nd > 1 ? (spread_bits16(q0) | (spread_bits16(quant(embd[(size_t)i * dim + 1], lo1, sc1)) << 1))
                    : q0;
}

// pos[order[q]] = q and deg[q] = row length of node order[q]
__global__ __launch_bounds__(kWg) void k_order_rows(const uint32_t* __restrict__ order,
                                                    const uint32_t* __restrict__ off, uint32_t n,
                                                    uint32_t* __restrict__ pos,
                                                    uint32_t* __restrict__ deg) {
    const uint32_t q = blockIdx.x * kWg + threadIdx.x;
    if (q >= n) return;
    const uint32_t i = order[q];
    pos[i] = q;
    deg[q] = off[i + 1] - off[i];
}

// row q of the renumbered CSR = row order[q] with its neighbours renamed
// through pos, entries in their order (kRadiiLanes lanes per row)
__global__ __launch_bounds__(kWg) void k_order_csr(const uint32_t* __restrict__ order,
                                                   const uint32_t* __restrict__ pos,
                                                   const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ other,
                                                   const uint32_t* __restrict__ doff, uint32_t n,
                                                   uint32_t* __restrict__ dcols) {
    const uint64_t t = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    const uint64_t q = t / kRadiiLanes;
    const uint32_t sub = (uint32_t)(t % kRadiiLanes);
    if (q >= n) return;
    const uint32_t i = order[q], b = off[i], e = off[i + 1], o = doff[q];
    for (uint32_t p = b + sub; p < e; p += kRadiiLanes) dcols[o + (p - b)] = pos[other[p]];
}

// pack[i] = (w, 1 / phi_sum, 1 / exp(re)) of node i, the per-node factors of
// the edge kernel's density term, read as one 16-byte row per endpoint
__global__ __launch_bounds__(kWg) void k_dens_pack(const float* __restrict__ w,
                                                   const float* __restrict__ phi,
                                                   const float* __restrict__ re, uint32_t n,
                                                   float4* __restrict__ pack) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    pack[i] = make_float4(w[i], 1.0f / phi[i], 1.0f / expf(re[i]), 0.0f);
}

// y[q] = embd[order[q]] (dim floats per row)
__global__ __launch_bounds__(kWg) void k_order_embd(const float* __restrict__ embd,
                                                    const uint32_t* __restrict__ order, uint32_t n,
                                                    int dim, float* __restrict__ y) {
    const uint32_t q = blockIdx.x * kWg + threadIdx.x;
    if (q >= n) return;
    const size_t src = (size_t)order[q] * dim, dst = (size_t)q * dim;
    for (int d = 0; d < dim; d++) y[dst + d] = embd[src + d];
}

__device__ inline double block_sum_f64(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kWg / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// Pass 1 of correlation_stats (src/utils/density.rs:471-481): per-block FP64
// partial sums of re over a fixed grid (so the order, and the result, is
// identical run to run).
__global__ __launch_bounds__(kWg) void k_dens_sum(const float* __restrict__ re, uint32_t n,
                                                  double* __restrict__ partials) {
    __shared__ double sh[kWg];
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * kWg + threadIdx.x; i < n; i += kDensBlocks * kWg)
        acc += (double)re[i];
    double s = block_sum_f64(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Pass 2 (:483-500): every block derives the same mean from the pass-1
// partials, then its centred partials of sum (re-mean)^2 and sum (re-mean)*r.
__global__ __launch_bounds__(kWg) void k_dens_centred(const float* __restrict__ re,
                                                      const float* __restrict__ r, uint32_t n,
                                                      double* __restrict__ partials) {
    __shared__ double sh[kWg];
    __shared__ double mean_sh;
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < kDensBlocks; q++) t += partials[q];
        mean_sh = t / (double)(n > 1 ? n : 1);
    }
    __syncthreads();
    const double mean = mean_sh;
    double ss = 0.0, cross = 0.0;
    for (uint32_t i = blockIdx.x * kWg + threadIdx.x; i < n; i += kDensBlocks * kWg) {
        double d = (double)re[i] - mean;
        ss += d * d;
        cross += d * (double)r[i];
    }
    double s1 = block_sum_f64(ss, sh);
    double s2 = block_sum_f64(cross, sh);
    if (threadIdx.x == 0) {
        partials[kDensBlocks + blockIdx.x] = s1;
        partials[2 * kDensBlocks + blockIdx.x] = s2;
    }
}

// Final reduction + umap_dens_weights (umap_optimiser_gpu.rs:568-586) + the
// dens_coeff of launch_dens_update (:1141-1142).  The f64 statistics are cast
// to f32 before use, as the reference's correlation_stats::<f32> returns them.
__global__ __launch_bounds__(kWg) void k_dens_weights(
    const float* __restrict__ re, const float* __restrict__ r, uint32_t n,
    const double* __restrict__ partials, float* __restrict__ weight,
    float* __restrict__ scalars, float var_shift, float lambda, float mu_tot) {
    __shared__ float st_sh[4];
    if (threadIdx.x == 0) {
        double t = 0.0, ss = 0.0, cross = 0.0;
        for (int q = 0; q < kDensBlocks; q++) t += partials[q];
        for (int q = 0; q < kDensBlocks; q++) ss += partials[kDensBlocks + q];
        for (int q = 0; q < kDensBlocks; q++) cross += partials[2 * kDensBlocks + q];
        double mean = t / (double)(n > 1 ? n : 1);
        double denom = (double)(n > 2 ? n - 1 : 1);
        double var = ss / denom, cov = cross / denom;
        double sd = sqrt(var + (double)var_shift);
        float mean_f = (float)mean, sd_f = (float)sd, cov_f = (float)cov;
        float denom_f = (float)(n > 2 ? n - 1 : 1);
        st_sh[0] = mean_f;
        st_sh[1] = sd_f;
        st_sh[2] = cov_f;
        st_sh[3] = lambda * mu_tot / (sd_f * denom_f);
        if (blockIdx.x == 0) {
            for (int q = 0; q < 4; q++) scalars[q] = st_sh[q];
        }
    }
    __syncthreads();
    const float mean_f = st_sh[0], sd_f = st_sh[1], cov_f = st_sh[2];
    const float inv_var = 1.0f / (sd_f * sd_f);
    for (uint32_t i = blockIdx.x * kWg + threadIdx.x; i < n; i += kDensBlocks * kWg)
        weight[i] = r[i] - cov_f * (re[i] - mean_f) * inv_var;
}

// ------------------------------------------------------ edge-parallel epoch

// One edge-parallel epoch: each wavefront owns kEdgeSpan consecutive edges
// and works through the firing ones compacted in LDS (epoch_edge_span).
// Built-in choice (the autotuner's fallback): spans of 128 edges from
// kWideSpanEdges edges (128 and 64 alike and 256 slower on the 1M fit graph:
// 317, 320 and 327 us), below that spans of 64, which leave enough
// wavefronts to fill the chip.
constexpr uint32_t kWideSpanEdges = 1u << 20;
constexpr int kWideSpan = 128;

// BITS: the firing edges come from the epoch's plane of the firing bitmask
// (k_fire_bits) instead of the schedule cursors, which that kernel advanced.
template <int DIM, bool DENS, int kEdgeSpan, bool BITS>
__global__ __launch_bounds__(kWg) void k_epoch_edges(
    const uint4* __restrict__ edges, const float* __restrict__ per, float* __restrict__ next,
    const float* __restrict__ embd, float* __restrict__ grad, uint32_t n_edges, uint32_t epoch,
    float epoch_f, uint32_t seed, FastMod fm, float a, float b, float two_a_b,
    float two_gamma_b, int neg_rate, const float* __restrict__ mu,
    const float4* __restrict__ dens_pack, const float* __restrict__ dens_scalars,
    const uint64_t* __restrict__ plane, uint64_t words, CompactRecords cr) {
    constexpr int kWaves = kWg / 64;
    __shared__ SpanQueue<kEdgeSpan> queue[kWaves];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const uint64_t base = ((uint64_t)blockIdx.x * kWaves + wave) * kEdgeSpan;
    uint32_t count;
    if constexpr (BITS)
        count = queue_from_bits<kEdgeSpan>(base, queue[wave], cr, plane, words);
    else
        count = queue_from_schedule<kEdgeSpan>(base, queue[wave], edges, per, next, n_edges,
                                               epoch_f);
    work_queue<DIM, DENS, kEdgeSpan>(queue[wave], count, embd, grad, epoch, seed, fm, a, b,
                                     two_a_b, two_gamma_b, neg_rate, mu, dens_pack, dens_scalars);
}

// The compact records of the bitmask epoch kernel (CompactRecords), from the
// 16-byte records: per edge {j, slot under j}; per 64-edge word the edges
// that start a run of equal i (ballot) and their count.  One wavefront per
// word.
__global__ __launch_bounds__(kWg) void k_run_starts(const uint4* __restrict__ edges,
                                                    uint32_t n_edges, uint64_t words,
                                                    uint2* __restrict__ pair,
                                                    uint64_t* __restrict__ run_start,
                                                    uint32_t* __restrict__ run_count) {
    const uint64_t e = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    const uint64_t w = e / 64;
    const int lane = threadIdx.x & 63;
    bool start = false;
    if (e < n_edges) {
        const uint4 r = edges[e];
        pair[e] = make_uint2(r.y, r.w);
        start = e == 0 || edges[e - 1].x != r.x;
    }
    const uint64_t m = __ballot(start);
    if (lane == 0 && w < words) {
        run_start[w] = m;
        run_count[w] = (uint32_t)__popcll(m);
    }
}

// run_info[run] = {i, slot under i of the run's first edge - its index} from
// each run's first edge; word_run[w] = (runs started before edge 64w) - 1
// (mod 2^32) from the exclusive scan of the words' run counts.
__global__ __launch_bounds__(kWg) void k_run_info(const uint4* __restrict__ edges,
                                                  uint32_t n_edges, uint64_t words,
                                                  const uint64_t* __restrict__ run_start,
                                                  const uint32_t* __restrict__ runs_before,
                                                  uint32_t* __restrict__ word_run,
                                                  uint2* __restrict__ run_info) {
    const uint64_t e = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    const uint64_t w = e / 64;
    const int lane = threadIdx.x & 63;
    if (w >= words) return;
    if (lane == 0) word_run[w] = runs_before[w] - 1u;
    if (e >= n_edges || !((run_start[w] >> lane) & 1ull)) return;
    const uint32_t run = runs_before[w] + (uint32_t)__popcll(run_start[w] & ((1ull << lane) - 1ull));
    const uint4 r = edges[e];
    run_info[run] = make_uint2(r.x, r.z - (uint32_t)e);
}

// Firing bitmask of a window of `count` epochs from `first`: each edge's
// schedule replayed exactly as the per-epoch scan advances it (the edge fires
// at epoch t when next <= t, then next += per, in f32), bit e mod 64 of word
// e / 64 of plane t - first set when edge e fires at epoch t (words past the
// last edge stay zero), and the cursors written back as they stand after the
// window.  A wavefront owns kFireWords consecutive words (64 edges each) and
// keeps their cursors and periods in registers through the window; per epoch
// it writes its words as one 128-byte row.  The epoch kernels then read one
// bit per edge instead of a cursor and a period (and a cursor written back).
constexpr int kFireWords = 16;
__global__ __launch_bounds__(kWg) void k_fire_bits(const float* __restrict__ per,
                                                   float* __restrict__ next, uint32_t n_edges,
                                                   uint64_t words, int32_t first, int32_t count,
                                                   uint64_t* __restrict__ bits) {
    const int lane = threadIdx.x & 63;
    const uint64_t w0 = (((uint64_t)blockIdx.x * kWg + threadIdx.x) >> 6) * kFireWords;
    if (w0 >= words) return;  // wavefront-uniform
    float nx[kFireWords], pr[kFireWords];
#pragma unroll
    for (int r = 0; r < kFireWords; r++) {
        const uint64_t e = (w0 + r) * 64 + lane;
        nx[r] = INFINITY;  // past the last edge: never fires
        pr[r] = 0.0f;
        if (e < n_edges) {
            nx[r] = next[e];
            pr[r] = per[e];
        }
    }
    for (int32_t t = 0; t < count; t++) {
        const float tf = (float)(first + t);
        uint64_t mine = 0;
#pragma unroll
        for (int r = 0; r < kFireWords; r++) {
            const bool fire = nx[r] <= tf;
            if (fire) nx[r] = nx[r] + pr[r];
            const uint64_t m = __ballot(fire);
            if (lane == r) mine = m;
        }
        if (lane < kFireWords && w0 + lane < words) bits[(size_t)t * words + w0 + lane] = mine;
    }
#pragma unroll
    for (int r = 0; r < kFireWords; r++) {
        const uint64_t e = (w0 + r) * 64 + lane;
        if (e < n_edges) next[e] = nx[r];
    }
}

template <int DIM>
__global__ __launch_bounds__(kWg) void k_adam_clear(float* __restrict__ grad,
                                                    float* __restrict__ m, float* __restrict__ v,
                                                    float* __restrict__ embd, uint32_t n,
                                                    float fused_lr, float epsc, float omb1,
                                                    float omb2, float* __restrict__ norm_partial) {
    const uint32_t node = blockIdx.x * kWg + threadIdx.x;
    if (node >= n) return;
    adam_clear_node<DIM>(node, grad, m, v, embd, fused_lr, epsc, omb1, omb2, norm_partial);
}

// grad = -0.0 everywhere: no node has an update yet (the edge strategy's
// cleared state; a kernel, not a memset, so it may sit in a captured graph)
__global__ __launch_bounds__(kWg) void k_fill_no_update(float* __restrict__ grad, uint64_t count) {
    const uint64_t q = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (q < count) grad[q] = -0.0f;
}

// has_update[i] from the -0.0 markers (what a caller reads after an edge epoch)
__global__ __launch_bounds__(kWg) void k_flags_from_marks(const float* __restrict__ grad,
                                                          uint32_t n, int dim,
                                                          uint32_t* __restrict__ flag) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    uint32_t any = 0;
    for (int d = 0; d < dim; d++) any |= __float_as_uint(grad[(size_t)i * dim + d]) != kNoUpdateBits;
    flag[i] = any;
}

// the -0.0 markers from has_update (a caller wrote gradients and flags)
__global__ __launch_bounds__(kWg) void k_marks_from_flags(float* __restrict__ grad,
                                                          const uint32_t* __restrict__ flag,
                                                          uint32_t n, int dim) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n || flag[i]) return;
    for (int d = 0; d < dim; d++) grad[(size_t)i * dim + d] = -0.0f;
}

template <class F>
void dispatch_dim(int dim, F&& f) {
    switch (dim) {
        case 1: f(Dim<1>{}); break;
        case 2: f(Dim<2>{}); break;
        case 3: f(Dim<3>{}); break;
        case 4: f(Dim<4>{}); break;
        default: f(Dim<0>{}); break;
    }
}

}  // namespace

// ----------------------------------------------------------- host side

AdamScalars adam_scalars(const mrs_umap_optim_params& p, int32_t epoch) {
    const float one = 1.0f;
    float lr_alpha = p.lr * (one - (float)epoch / (float)p.n_epochs);
    float t = (float)(epoch + 1);
    float beta1t = powf(p.beta1, t);
    float beta2t = powf(p.beta2, t);
    float sqrt_b2t1 = sqrtf(one - beta2t);
    float ad_scale = sqrt_b2t1 / (one - beta1t);
    AdamScalars s;
    s.epsc = sqrt_b2t1 * p.eps;
    s.fused_lr = lr_alpha * ad_scale;
    return s;
}

bool dens_is_active(const mrs_dens_params& dp, int32_t epoch, int32_t n_epochs) {
    if (dp.lambda <= 0.0f || n_epochs == 0) return false;
    float progress = (float)(epoch + 1) / (float)n_epochs;
    return progress > 1.0f - dp.frac;
}

void launch_grad_accum_node(mrs_umap_state* st, int32_t epoch, uint32_t seed, bool dens) {
    const mrs_umap_optim_params& p = st->params;
    const float two = 2.0f;
    const uint32_t n = (uint32_t)st->n;
    FastMod fm = make_fastmod(n);
    hipStream_t s = st->ctx->stream;
    dispatch_dim(st->dim, [&](auto dd) {
        constexpr int D = decltype(dd)::value;
        auto kern = dens ? k_grad_accum_node<D, true> : k_grad_accum_node<D, false>;
        hipLaunchKernelGGL(kern, dim3(div_up(n, kWg)), dim3(kWg), 0, s, st->embd.p, st->off.p,
                           st->csr_edge_idx.p, st->csr_other.p, st->next.p, st->grad.p,
                           st->has_update.p, n, st->dim, (uint32_t)epoch, (float)epoch, seed, fm,
                           p.a, p.b, two * p.a * p.b, two * p.gamma * p.b, p.neg_sample_rate,
                           st->mu.p, st->dens_re.p, st->dens_phi.p, st->dens_w.p,
                           st->dens_scalars.p);
    });
    MRS_CHECK_LAUNCH();
    st->grad_clean = false;
}

void launch_adam_elem(mrs_umap_state* st, int32_t epoch) {
    AdamScalars sc = adam_scalars(st->params, epoch);
    uint64_t total = (uint64_t)st->n * st->dim;
    hipLaunchKernelGGL(k_adam_elem, dim3(div_up(total, kWg)), dim3(kWg), 0, st->ctx->stream,
                       st->grad.p, st->has_update.p, st->m.p, st->v.p, st->embd.p, total,
                       st->dim, sc.fused_lr, sc.epsc, 1.0f - st->params.beta1,
                       1.0f - st->params.beta2);
    MRS_CHECK_LAUNCH();
}

void launch_schedule(mrs_umap_state* st, int32_t epoch) {
    if (st->n_edges == 0) return;
    hipLaunchKernelGGL(k_schedule, dim3(div_up(st->n_edges, kWg)), dim3(kWg), 0,
                       st->ctx->stream, st->per.p, st->next.p, (uint64_t)st->n_edges,
                       (float)epoch);
    MRS_CHECK_LAUNCH();
}

void launch_dens_update_dev(mrs_umap_state* st, bool reference_order) {
    if (!st->dens_enabled) throw Error(MRS_INVALID_ARGUMENT, "state has no densMAP buffers");
    const uint32_t n = (uint32_t)st->n;
    hipStream_t s = st->ctx->stream;
    dispatch_dim(st->dim, [&](auto dd) {
        constexpr int D = decltype(dd)::value;
        if (D > 0 && !reference_order) {
            constexpr int DL = D > 0 ? D : 1;
            if (!st->dens_order.p) {
                hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
                MRS_HIP(hipStreamIsCapturing(s, &cap));
                if (cap == hipStreamCaptureStatusNone) state_dens_order(st);
            }
            const bool ordered = st->dens_order.p != nullptr;
            if (ordered) {
                hipLaunchKernelGGL(k_order_embd, dim3(div_up(n, kWg)), dim3(kWg), 0, s, st->embd.p,
                                   st->dens_order.p, n, st->dim, st->dens_y.p);
            }
            hipLaunchKernelGGL(k_dens_radii_lanes<DL>, dim3(div_up((uint64_t)n * kRadiiLanes, kWg)),
                               dim3(kWg), 0, s, ordered ? st->dens_y.p : st->embd.p,
                               ordered ? st->dens_off.p : st->off.p,
                               ordered ? st->dens_cols.p : st->csr_other.p,
                               ordered ? st->dens_order.p : nullptr, st->dens_re.p,
                               st->dens_phi.p, n, st->params.a, st->params.b, (float)kDensEps,
                               FLT_EPSILON);
        } else {
            hipLaunchKernelGGL(k_dens_radii<D>, dim3(div_up(n, kWg)), dim3(kWg), 0, s, st->embd.p,
                               st->off.p, st->csr_other.p, st->dens_re.p, st->dens_phi.p, n,
                               st->dim, st->params.a, st->params.b, (float)kDensEps, FLT_EPSILON);
        }
    });
    MRS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_dens_sum, dim3(kDensBlocks), dim3(kWg), 0, s, st->dens_re.p, n,
                       st->dens_partials.p);
    MRS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_dens_centred, dim3(kDensBlocks), dim3(kWg), 0, s, st->dens_re.p,
                       st->dens_r.p, n, st->dens_partials.p);
    MRS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_dens_weights, dim3(kDensBlocks), dim3(kWg), 0, s, st->dens_re.p,
                       st->dens_r.p, n, st->dens_partials.p, st->dens_w.p, st->dens_scalars.p,
                       st->dens_params.var_shift, st->dens_params.lambda, st->mu_tot);
    MRS_CHECK_LAUNCH();
    dens_pack_refresh(st);
    MRS_CHECK_LAUNCH();
}

// The edge strategy's cleared state: every gradient at the -0.0 marker.
void clear_update_marks(mrs_umap_state* st) {
    const uint64_t count = (uint64_t)st->n * st->dim;
    hipLaunchKernelGGL(k_fill_no_update, dim3(div_up(count, kWg)), dim3(kWg), 0, st->ctx->stream,
                       st->grad.p, count);
    MRS_CHECK_LAUNCH();
    st->grad_clean = true;
}

void derive_update_flags(mrs_umap_state* st) {
    const uint32_t n = (uint32_t)st->n;
    hipLaunchKernelGGL(k_flags_from_marks, dim3(div_up(n, kWg)), dim3(kWg), 0, st->ctx->stream,
                       st->grad.p, n, st->dim, st->has_update.p);
    MRS_CHECK_LAUNCH();
}

void marks_from_flags(mrs_umap_state* st) {
    const uint32_t n = (uint32_t)st->n;
    hipLaunchKernelGGL(k_marks_from_flags, dim3(div_up(n, kWg)), dim3(kWg), 0, st->ctx->stream,
                       st->grad.p, st->has_update.p, n, st->dim);
    MRS_CHECK_LAUNCH();
}

namespace {

// shape class of an edge epoch: dimension, density term, and edges and nodes
// as powers of two (the node count sets how much of the embedding, the target
// of the random gathers, each XCD's L2 holds)
uint64_t epoch_span_shape(const mrs_umap_state* st, bool dens) {
    return (uint64_t)st->dim | (uint64_t)dens << 4 | (uint64_t)ceil_log2(st->n_edges) << 8 |
           (uint64_t)ceil_log2(st->n) << 16;
}

}  // namespace

int epoch_span_choice(const mrs_umap_state* st, bool dens) {
    const int forced = tune_forced(kTuneEpochSpan);
    if (forced > 0) return forced;
    const int rec = tune_lookup(kTuneEpochSpan, epoch_span_shape(st, dens));
    if (rec > 0) return rec;
    return (uint32_t)st->n_edges >= kWideSpanEdges ? kWideSpan : 64;
}

void launch_epoch_edges(mrs_umap_state* st, int32_t epoch, uint32_t seed, bool dens, int span,
                        const uint64_t* plane) {
    if (!st->grad_clean) clear_update_marks(st);
    const mrs_umap_optim_params& p = st->params;
    const float two = 2.0f;
    FastMod fm = make_fastmod((uint32_t)st->n);
    const uint32_t ne = (uint32_t)st->n_edges;
    const uint64_t words = div_up((uint64_t)ne, (uint64_t)64);
    hipStream_t s = st->ctx->stream;
    if (plane && !st->rec_pair.p) throw Error(MRS_DEVICE_ERROR, "compact edge records not built");
    const CompactRecords cr{st->rec_pair.p, st->run_start.p, st->word_run.p, st->run_info.p};
    if (span == 0) span = epoch_span_choice(st, dens);
    if (span != 64 && span != 128 && span != 256)
        throw Error(MRS_INVALID_ARGUMENT, "edge epoch span must be 64, 128 or 256");
    // one wavefront per span of edges
    const unsigned blocks = div_up(div_up(ne, (unsigned)span), kWg / 64);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kWg), 0, s, st->edges.p, st->per.p,
                           st->next.p, st->embd.p, st->grad.p, ne, (uint32_t)epoch, (float)epoch,
                           seed, fm, p.a, p.b, two * p.a * p.b, two * p.gamma * p.b,
                           p.neg_sample_rate, st->mu.p,
                           reinterpret_cast<const float4*>(st->dens_pack.p), st->dens_scalars.p,
                           plane, words, cr);
    };
    auto by_span = [&](auto d) {
        constexpr int D = decltype(d)::value;
        auto by_source = [&](auto sp) {
            constexpr int S = decltype(sp)::value;
            if (plane) dens ? go(k_epoch_edges<D, true, S, true>) : go(k_epoch_edges<D, false, S, true>);
            else dens ? go(k_epoch_edges<D, true, S, false>) : go(k_epoch_edges<D, false, S, false>);
        };
        if (span == 256) by_source(Dim<256>{});
        else if (span == 128) by_source(Dim<128>{});
        else by_source(Dim<64>{});
    };
    switch (st->dim) {
        case 1: by_span(Dim<1>{}); break;
        case 2: by_span(Dim<2>{}); break;
        case 3: by_span(Dim<3>{}); break;
        case 4: by_span(Dim<4>{}); break;
        case 8: by_span(Dim<8>{}); break;
        default:
            throw Error(MRS_INVALID_ARGUMENT,
                        "edge-parallel epoch supports n_dim in {1,2,3,4,8}; use node-exact");
    }
    MRS_CHECK_LAUNCH();
}

void build_compact_records(mrs_umap_state* st) {
    const uint32_t ne = (uint32_t)st->n_edges;
    const uint64_t words = div_up((uint64_t)ne, (uint64_t)64);
    if (!words) return;
    hipStream_t s = st->ctx->stream;
    st->rec_pair.alloc(ne);
    st->run_start.alloc(words);
    st->word_run.alloc(words);
    st->run_info.alloc((size_t)std::min<int64_t>(st->n, st->n_edges));
    DevBuf<uint32_t> runs_before(words);
    const unsigned blocks = div_up(words * 64, kWg);
    hipLaunchKernelGGL(k_run_starts, dim3(blocks), dim3(kWg), 0, s, st->edges.p, ne, words,
                       st->rec_pair.p, st->run_start.p, runs_before.p);
    MRS_CHECK_LAUNCH();
    exclusive_scan_u32(runs_before.p, runs_before.p, words, nullptr, s);
    hipLaunchKernelGGL(k_run_info, dim3(blocks), dim3(kWg), 0, s, st->edges.p, ne, words,
                       st->run_start.p, runs_before.p, st->word_run.p, st->run_info.p);
    MRS_CHECK_LAUNCH();
    // runs_before returns to the pool at scope exit; the stream orders its reuse
}

void launch_fire_bits(mrs_umap_state* st, int32_t first, int32_t count) {
    const uint32_t ne = (uint32_t)st->n_edges;
    const uint64_t words = div_up((uint64_t)ne, (uint64_t)64);
    if (!words || count <= 0) return;
    if (count > kFireWindow || st->fire_bits.n < (size_t)count * words)
        throw Error(MRS_DEVICE_ERROR, "firing bitmask smaller than the window");
    const uint64_t waves = div_up(words, (uint64_t)kFireWords);
    hipLaunchKernelGGL(k_fire_bits, dim3((unsigned)div_up(waves, (uint64_t)(kWg / 64))), dim3(kWg),
                       0, st->ctx->stream, st->per.p, st->next.p, ne, words, first, count,
                       st->fire_bits.p);
    MRS_CHECK_LAUNCH();
}

void launch_adam_clear(mrs_umap_state* st, int32_t epoch, float* norm_partial) {
    AdamScalars sc = adam_scalars(st->params, epoch);
    const uint32_t n = (uint32_t)st->n;
    const float omb1 = 1.0f - st->params.beta1, omb2 = 1.0f - st->params.beta2;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(div_up(n, kWg)), dim3(kWg), 0, st->ctx->stream,
                           st->grad.p, st->m.p, st->v.p, st->embd.p, n, sc.fused_lr, sc.epsc,
                           omb1, omb2, norm_partial);
    };
    switch (st->dim) {
        case 1: go(k_adam_clear<1>); break;
        case 2: go(k_adam_clear<2>); break;
        case 3: go(k_adam_clear<3>); break;
        case 4: go(k_adam_clear<4>); break;
        case 8: go(k_adam_clear<8>); break;
        default: throw Error(MRS_INVALID_ARGUMENT, "unsupported n_dim for adam_clear");
    }
    MRS_CHECK_LAUNCH();
}

namespace {

// RAII bracket that records a start/stop event pair around one launch when the
// state is profiling.
struct ProfScope {
    mrs_umap_state* st;
    int kind;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(mrs_umap_state* s, int k) : st(s), kind(k) {
        if (!st->profiling) return;
        MRS_HIP(hipEventCreate(&a));
        MRS_HIP(hipEventCreate(&b));
        MRS_HIP(hipEventRecord(a, st->ctx->stream));
    }
    ~ProfScope() {
        if (!st->profiling) return;
        (void)hipEventRecord(b, st->ctx->stream);
        st->prof_events.push_back({kind, {a, b}});
    }
};

}  // namespace

void prof_collect(mrs_umap_state* st) {
    for (auto& ev : st->prof_events) {
        float ms = 0.0f;
        MRS_HIP(hipEventSynchronize(ev.second.second));
        MRS_HIP(hipEventElapsedTime(&ms, ev.second.first, ev.second.second));
        st->prof_ms[ev.first] += ms;
        st->prof_count[ev.first] += 1;
        (void)hipEventDestroy(ev.second.first);
        (void)hipEventDestroy(ev.second.second);
    }
    st->prof_events.clear();
}

void run_epochs(mrs_umap_state* st, int32_t first, int32_t count, uint32_t seed, int strategy) {
    const int32_t N = st->params.n_epochs;
    if (strategy != MRS_EPOCH_EDGE_ATOMIC && strategy != MRS_EPOCH_NODE_EXACT)
        throw Error(MRS_INVALID_ARGUMENT, "unknown epoch strategy");
    if (strategy == MRS_EPOCH_EDGE_ATOMIC) {
        int d = st->dim;
        if (!(d == 1 || d == 2 || d == 3 || d == 4 || d == 8)) strategy = MRS_EPOCH_NODE_EXACT;
    }
    // Edge-parallel runs of kFireMinEpochs or more take their firing edges
    // from bitmask windows (launch_fire_bits).  The buffer is allocated here,
    // before any capture; a captured loop holds its address, so a new buffer
    // drops the captured loop.
    const uint64_t fire_words = div_up((uint64_t)st->n_edges, (uint64_t)64);
    const bool use_bits =
        strategy == MRS_EPOCH_EDGE_ATOMIC && count >= kFireMinEpochs && fire_words > 0;
    if (use_bits) {
        const size_t need = (size_t)std::min(count, kFireWindow) * fire_words;
        if (st->fire_bits.n < need) {
            if (st->graph_exec) {
                MRS_HIP(hipStreamSynchronize(st->ctx->stream));
                (void)hipGraphExecDestroy(st->graph_exec);
                st->graph_exec = nullptr;
                std::fill(st->graph_key, st->graph_key + 4, -1);
            }
            StreamScope scope(st->ctx->stream, st->ctx->pool);
            st->fire_bits.alloc(need);
        }
        if (!st->rec_pair.p) {
            StreamScope scope(st->ctx->stream, st->ctx->pool);
            build_compact_records(st);
        }
    }
    auto plane_of = [&](int32_t e) -> const uint64_t* {
        return use_bits ? st->fire_bits.p + (size_t)((e - first) % kFireWindow) * fire_words
                        : nullptr;
    };
    // Span tuning (autotune.hpp).  For each stretch of the run with one
    // density flag whose shape class is undecided, a window of kTuneEpochs
    // epochs alternates the candidate spans with an event pair around each
    // edge launch, and the median per span decides.  The window sits in the
    // middle of the stretch: in the first epochs only the edges whose period
    // has already come round fire, so they are lighter than the rest.
    // Directly launched runs only (not under capture).
    constexpr int kSpans[3] = {64, 128, 256};
    constexpr int kTuneEpochs = 12;
    struct Window {
        int32_t start;
        bool dens;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    };
    std::vector<Window> windows;
    auto drop_windows = [&]() {
        for (auto& w : windows)
            for (auto& ev : w.ev) {
                if (ev.first) (void)hipEventDestroy(ev.first);
                if (ev.second) (void)hipEventDestroy(ev.second);
            }
        windows.clear();
    };
    auto plan_tuning = [&]() {
        if (strategy != MRS_EPOCH_EDGE_ATOMIC || !tune_enabled() || tune_forced(kTuneEpochSpan) > 0)
            return;
        auto dens_at = [&](int32_t e) {
            return st->dens_enabled && dens_is_active(st->dens_params, e, N);
        };
        for (int32_t a = first; a < first + count;) {
            const bool d = dens_at(a);
            int32_t b = a + 1;
            while (b < first + count && dens_at(b) == d) b++;
            const int32_t len = b - a;
            bool planned = false;
            for (auto& w : windows) planned |= w.dens == d;
            if (len >= kTuneEpochs && !planned &&
                tune_lookup(kTuneEpochSpan, epoch_span_shape(st, d)) <= 0) {
                Window w{a + (len - kTuneEpochs) / 2, d, {}};
                w.ev.resize(kTuneEpochs, {nullptr, nullptr});
                windows.push_back(std::move(w));
                for (auto& ev : windows.back().ev) {
                    MRS_HIP(hipEventCreate(&ev.first));
                    MRS_HIP(hipEventCreate(&ev.second));
                }
            }
            a = b;
        }
    };
    auto finish_window = [&](Window& w) {
        std::vector<std::vector<float>> t(3);
        for (int q = 0; q < kTuneEpochs; q++) {
            float ms = 0.0f;
            MRS_HIP(hipEventSynchronize(w.ev[q].second));
            MRS_HIP(hipEventElapsedTime(&ms, w.ev[q].first, w.ev[q].second));
            t[q % 3].push_back(ms);
        }
        std::vector<float> med;
        for (auto& v : t) {
            std::sort(v.begin(), v.end());
            med.push_back(v[v.size() / 2]);
        }
        tune_record(kTuneEpochSpan, epoch_span_shape(st, w.dens), {kSpans[0], kSpans[1], kSpans[2]},
                    med);
    };
    auto window_of = [&](int32_t e) -> Window* {
        for (auto& w : windows)
            if (e >= w.start && e < w.start + kTuneEpochs) return &w;
        return nullptr;
    };
    auto launches = [&]() {
        for (int32_t e = first; e < first + count; e++) {
            if (use_bits && (e - first) % kFireWindow == 0) {
                ProfScope ps(st, 6);
                launch_fire_bits(st, e, std::min(kFireWindow, first + count - e));
            }
            bool dens = st->dens_enabled && dens_is_active(st->dens_params, e, N);
            if (dens) {
                ProfScope ps(st, 5);
                launch_dens_update_dev(st, strategy == MRS_EPOCH_NODE_EXACT);
            }
            if (strategy == MRS_EPOCH_NODE_EXACT) {
                { ProfScope ps(st, 2); launch_grad_accum_node(st, e, seed, dens); }
                { ProfScope ps(st, 3); launch_adam_elem(st, e); }
                { ProfScope ps(st, 4); launch_schedule(st, e); }
            } else if (Window* w = window_of(e)) {
                const int q = e - w->start;
                if (!st->grad_clean) clear_update_marks(st);
                MRS_HIP(hipEventRecord(w->ev[q].first, st->ctx->stream));
                {
                    ProfScope ps(st, 0);
                    launch_epoch_edges(st, e, seed, dens, kSpans[q % 3], plane_of(e));
                }
                MRS_HIP(hipEventRecord(w->ev[q].second, st->ctx->stream));
                { ProfScope ps(st, 1); launch_adam_clear(st, e, nullptr); }
                if (q + 1 == kTuneEpochs) finish_window(*w);
            } else {
                { ProfScope ps(st, 0); launch_epoch_edges(st, e, seed, dens, 0, plane_of(e)); }
                { ProfScope ps(st, 1); launch_adam_clear(st, e, nullptr); }
            }
        }
    };
    hipStream_t s = st->ctx->stream;
    if (st->profiling || !st->use_graphs || count < 16) {
        try {
            plan_tuning();
            launches();
        } catch (...) {
            drop_windows();
            throw;
        }
        drop_windows();
        return;
    }
    // Capture the whole loop once and replay it: one launch per fit instead of
    // 2-3 per epoch.  The accumulate-by-atomics buffer is cleared outside the
    // capture.
    if (strategy == MRS_EPOCH_EDGE_ATOMIC && !st->grad_clean) clear_update_marks(st);
    const int64_t key[4] = {first, count, (int64_t)seed, strategy};
    const bool hit = st->graph_exec && std::equal(key, key + 4, st->graph_key);
    if (!hit && !std::equal(key, key + 4, st->last_key)) {
        std::copy(key, key + 4, st->last_key);
        try {
            plan_tuning();
            launches();
        } catch (...) {
            drop_windows();
            throw;
        }
        drop_windows();
        return;
    }
    if (!hit) {
        if (st->graph_exec) {
            (void)hipGraphExecDestroy(st->graph_exec);
            st->graph_exec = nullptr;
        }
        MRS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            launches();
        } catch (...) {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(s, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            throw;
        }
        hipGraph_t graph = nullptr;
        MRS_HIP(hipStreamEndCapture(s, &graph));
        hipError_t ie = hipGraphInstantiate(&st->graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        MRS_HIP(ie);
        std::copy(key, key + 4, st->graph_key);
    }
    MRS_HIP(hipGraphLaunch(st->graph_exec, s));
    if (strategy == MRS_EPOCH_NODE_EXACT) st->grad_clean = false;
}

// ------------------------------------------------------------- state build

HostCsr host_csr_from_graph(int64_t n, const int64_t* adj_off, const int32_t* adj_cols,
                            const float* adj_vals) {
    if (n == 0) throw Error(MRS_NO_DATA, "UMAP: empty data (NoData)");
    HostCsr c;
    c.n = n;
    std::vector<uint32_t> degree((size_t)n, 0);
    int64_t ne = 0;
    float max_w = 0.0f;
    for (int64_t i = 0; i < n; i++)
        for (int64_t p = adj_off[i]; p < adj_off[i + 1]; p++) {
            int64_t j = adj_cols[p];
            if (j < 0 || j >= n) throw Error(MRS_INVALID_ARGUMENT, "graph column out of range");
            if (i < j) {
                degree[i]++;
                degree[j]++;
                if (adj_vals[p] > max_w) max_w = adj_vals[p];
                ne++;
            }
        }
    if (ne == 0) throw Error(MRS_NO_GRAPH_EDGES, "UMAP: no edges to optimise (NoGraphEdges)");
    if (2 * ne > (int64_t)UINT32_MAX) throw Error(MRS_INVALID_ARGUMENT, "graph too large");
    c.n_edges = ne;
    c.off.assign((size_t)n + 1, 0);
    for (int64_t i = 0; i < n; i++) c.off[i + 1] = c.off[i] + degree[i];
    c.csr_edge_idx.resize((size_t)(2 * ne));
    c.csr_other.resize((size_t)(2 * ne));
    c.edges.resize((size_t)ne);
    c.per.resize((size_t)ne);
    c.mu.resize((size_t)ne);
    std::vector<uint32_t> cursor(c.off.begin(), c.off.end() - 1);
    int64_t e = 0;
    for (int64_t i = 0; i < n; i++)
        for (int64_t p = adj_off[i]; p < adj_off[i + 1]; p++) {
            int64_t j = adj_cols[p];
            if (!(i < j)) continue;
            uint32_t pi = cursor[i]++, pj = cursor[j]++;
            c.csr_edge_idx[pi] = (uint32_t)e;
            c.csr_other[pi] = (uint32_t)j;
            c.csr_edge_idx[pj] = (uint32_t)e;
            c.csr_other[pj] = (uint32_t)i;
            c.edges[e] = make_uint4((uint32_t)i, (uint32_t)j, pi - c.off[i], pj - c.off[j]);
            float w = adj_vals[p];
            float norm = w / max_w;
            c.per[e] = norm > 0.0f ? 1.0f / norm : 1e8f;
            c.mu[e] = w;
            e++;
        }
    return c;
}

void state_alloc_nodes(mrs_umap_state* st, const float* embd) {
    hipStream_t s = st->ctx->stream;
    const size_t nd = (size_t)st->n * st->dim;
    st->embd.alloc(nd);
    st->m.alloc(nd);
    st->v.alloc(nd);
    st->grad.alloc(nd);
    st->has_update.alloc((size_t)st->n);
    st->norm_partial.alloc((size_t)st->n);
    MRS_HIP(hipMemcpyAsync(st->embd.p, embd, nd * sizeof(float), hipMemcpyDefault, s));
    MRS_HIP(hipMemsetAsync(st->m.p, 0, st->m.bytes(), s));
    MRS_HIP(hipMemsetAsync(st->v.p, 0, st->v.bytes(), s));
    MRS_HIP(hipMemsetAsync(st->grad.p, 0, st->grad.bytes(), s));
    MRS_HIP(hipMemsetAsync(st->has_update.p, 0, st->has_update.bytes(), s));
    st->grad_clean = false;  // zeros, not the edge strategy's -0.0 markers
    st->dens_scalars.alloc(4);
    MRS_HIP(hipMemsetAsync(st->dens_scalars.p, 0, st->dens_scalars.bytes(), s));
    if (!st->dens_enabled) {
        // length-1 dummies keep every kernel argument valid on the plain path
        st->dens_r.alloc(1);
        st->dens_re.alloc(1);
        st->dens_phi.alloc(1);
        st->dens_w.alloc(1);
        st->dens_pack.alloc(4);
    }
}

// Rebuild the packed per-node density factors from dens_w, dens_phi and
// dens_re (after each density update, and after a caller writes one of them).
void dens_pack_refresh(mrs_umap_state* st) {
    if (!st->dens_enabled) return;
    const uint32_t n = (uint32_t)st->n;
    hipLaunchKernelGGL(k_dens_pack, dim3(div_up(n, kWg)), dim3(kWg), 0, st->ctx->stream,
                       st->dens_w.p, st->dens_phi.p, st->dens_re.p, n,
                       reinterpret_cast<float4*>(st->dens_pack.p));
    MRS_CHECK_LAUNCH();
}

// The densMAP locality order of a state (see k_morton_keys), for the
// dimensions the lane-parallel radii kernel serves (1..4).  It is built from
// the embedding at the state's first density update outside a stream capture
// (by then the layout has pulled graph neighbours together), and kept.
void state_dens_order(mrs_umap_state* st) {
    hipStream_t s = st->ctx->stream;
    const uint32_t n = (uint32_t)st->n;
    const int dim = st->dim;
    if (n < 2 || dim < 1 || dim > 4) return;
    const int nd = dim >= 2 ? 2 : 1;
    DevBuf<uint32_t> bbox(4), key(n), skey(n), pos(n), deg((size_t)n + 1);
    const uint32_t init[4] = {~0u, 0u, ~0u, 0u};
    MRS_HIP(hipMemcpyAsync(bbox.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_embd_bbox, dim3(std::min<uint32_t>(div_up(n, kWg), 1024)), dim3(kWg), 0,
                       s, st->embd.p, n, dim, nd, bbox.p);
    MRS_CHECK_LAUNCH();
    uint32_t bb[4];
    MRS_HIP(hipMemcpyAsync(bb, bbox.p, sizeof(bb), hipMemcpyDeviceToHost, s));
    MRS_HIP(hipStreamSynchronize(s));
    auto to_float = [](uint32_t o) {
        const uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
        float f;
        std::memcpy(&f, &u, sizeof(f));
        return f;
    };
    float lo[2] = {0.0f, 0.0f}, sc[2] = {0.0f, 0.0f};
    for (int d = 0; d < nd; d++) {
        lo[d] = to_float(bb[2 * d]);
        const float span = to_float(bb[2 * d + 1]) - lo[d];
        sc[d] = std::isfinite(span) && span > 0.0f ? 65535.0f / span : 0.0f;
        if (!std::isfinite(lo[d])) lo[d] = 0.0f;
    }
    hipLaunchKernelGGL(k_morton_keys, dim3(div_up(n, kWg)), dim3(kWg), 0, s, st->embd.p, n, dim,
                       nd, lo[0], sc[0], lo[1], sc[1], key.p);
    MRS_CHECK_LAUNCH();
    st->dens_order.alloc(n);
    radix_sort_pairs_u32(key.p, nullptr, n, 32, skey.p, st->dens_order.p, s);
    hipLaunchKernelGGL(k_order_rows, dim3(div_up(n, kWg)), dim3(kWg), 0, s, st->dens_order.p,
                       st->off.p, n, pos.p, deg.p);
    MRS_CHECK_LAUNCH();
    st->dens_off.alloc((size_t)n + 1);
    exclusive_scan_u32(deg.p, st->dens_off.p, (uint64_t)n + 1, nullptr, s);
    const size_t nnz = 2 * (size_t)st->n_edges;
    st->dens_cols.alloc(nnz ? nnz : 1);
    hipLaunchKernelGGL(k_order_csr, dim3(div_up((uint64_t)n * kRadiiLanes, kWg)), dim3(kWg), 0, s,
                       st->dens_order.p, pos.p, st->off.p, st->csr_other.p, st->dens_off.p, n,
                       st->dens_cols.p);
    MRS_CHECK_LAUNCH();
    st->dens_y.alloc((size_t)n * dim);
    MRS_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
}

void state_alloc_dens(mrs_umap_state* st, const mrs_dens_params& dp, const float* r,
                      float mu_tot) {
    hipStream_t s = st->ctx->stream;
    const size_t n = (size_t)st->n;
    st->dens_enabled = true;
    st->dens_params = dp;
    st->mu_tot = mu_tot;
    st->dens_r.alloc(n);
    st->dens_re.alloc(n);
    st->dens_phi.alloc(n);
    st->dens_w.alloc(n);
    st->dens_partials.alloc(3 * kDensBlocks);
    st->dens_pack.alloc(4 * n);
    MRS_HIP(hipMemsetAsync(st->dens_pack.p, 0, st->dens_pack.bytes(), s));
    MRS_HIP(hipMemcpyAsync(st->dens_r.p, r, n * sizeof(float), hipMemcpyDefault, s));
    MRS_HIP(hipMemsetAsync(st->dens_re.p, 0, st->dens_re.bytes(), s));
    MRS_HIP(hipMemsetAsync(st->dens_phi.p, 0, st->dens_phi.bytes(), s));
    MRS_HIP(hipMemsetAsync(st->dens_w.p, 0, st->dens_w.bytes(), s));
}

}  // namespace mrs

mrs_umap_state::~mrs_umap_state() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    for (auto& ev : prof_events) {
        (void)hipEventDestroy(ev.second.first);
        (void)hipEventDestroy(ev.second.second);
    }
}

namespace mrs {

mrs_umap_state* make_state(mrs_ctx* ctx, const float* embd, int64_t n, int32_t dim,
                           const HostCsr& c, const mrs_umap_optim_params& p,
                           const mrs_dens_state* dens) {
    auto* st = new mrs_umap_state();
    try {
        st->ctx = ctx;
        st->n = n;
        st->dim = dim;
        st->n_edges = c.n_edges;
        st->params = p;
        hipStream_t s = ctx->stream;
        auto up = [&](auto& buf, const auto& vec) {
            buf.alloc(vec.size());
            MRS_HIP(hipMemcpyAsync(buf.p, vec.data(), buf.bytes(), hipMemcpyHostToDevice, s));
        };
        up(st->off, c.off);
        up(st->csr_edge_idx, c.csr_edge_idx);
        up(st->csr_other, c.csr_other);
        up(st->edges, c.edges);
        up(st->per, c.per);
        up(st->next, c.per);  // cursors start one period from epoch 0
        up(st->mu, c.mu);
        if (dens) state_alloc_dens(st, dens->params, dens->r, dens->mu_tot);
        state_alloc_nodes(st, embd);
        MRS_HIP(hipStreamSynchronize(s));  // host vectors die on return
    } catch (...) {
        delete st;
        throw;
    }
    return st;
}

}  // namespace mrs
