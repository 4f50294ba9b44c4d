// Spectral initial layout on the device: spectral_layout (src/data/init.rs:986-1004).
//
//   find_connected_components (:349-391)  min-label hooking + pointer jumping;
//       a component's id is its smallest vertex, which orders components the
//       way the reference's BFS from ascending start vertices does.
//   graph_to_negative_normalised_adjacency (:210-256)  M = -D^-1/2 W D^-1/2
//       (degrees accumulated per row in COO order, f64).
//   compute_smallest_eigenpairs_lanczos (src/utils/math.rs:401-627)  Lanczos
//       with full re-orthogonalisation in f64 (the reference: modified
//       Gram-Schmidt; here classical Gram-Schmidt applied twice, equally
//       orthogonal, a few blocked launches per step instead of two per basis
//       vector), run for ALL components at once: vectors are global, the SpMV
//       (16 lanes per row over precomputed f64 matrix values) is block
//       diagonal by construction, and every inner product is a segmented
//       reduction over the component's vertices (fixed chunking, so the result
//       is identical run to run).  The small tridiagonal problems are solved
//       on the host (cyclic Jacobi).
//   finalise_spectral_embedding (:905-960) / multi_component_init (:413-537)
//       with the simplex / data-centroid spectral / scatter meta layouts
//       (:539-766).
//
// The reference's StdRng streams (Lanczos start vector, noise, random fills)
// cannot be reproduced here; counter-based hashes draw from the same
// distributions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <exception>
#include <numeric>
#include <thread>
#include <vector>

#include "spectral.hpp"

namespace mrs {

namespace {

constexpr int kWg = 256;
constexpr int kChunk = 1024;      // rows per reduction chunk (fixed: results are run-to-run identical)
constexpr int kCenChunk = 256;    // rows per centroid chunk

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline double u01(uint64_t seed, uint64_t stream, uint64_t i) {
    uint64_t z = mix64(seed * 0x9E3779B97F4A7C15ull + stream * 0xD1B54A32D192ED03ull + i);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__global__ __launch_bounds__(kWg) void k_row_offsets(const uint32_t* __restrict__ coo_row,
                                                     uint64_t nnz, uint32_t n,
                                                     uint32_t* __restrict__ off) {
    uint32_t v = blockIdx.x * kWg + threadIdx.x;
    if (v > n) return;
    uint64_t lo = 0, hi = nnz;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (coo_row[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    off[v] = (uint32_t)lo;
}

__global__ __launch_bounds__(kWg) void k_degrees(const uint32_t* __restrict__ off,
                                                 const float* __restrict__ val, uint32_t n,
                                                 double* __restrict__ dinv) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    double dg = 0.0;
    for (uint32_t p = off[i]; p < off[i + 1]; p++) dg += (double)val[p];
    dinv[i] = dg > 1e-8 ? 1.0 / sqrt(dg) : 0.0;
}

__global__ __launch_bounds__(kWg) void k_cc_init(uint32_t* __restrict__ lab, uint32_t n) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i < n) lab[i] = i;
}

// Hook the components of u and v together: the larger of their labels is
// lowered to the smaller one (a root's label is itself after a jump pass).
// Labels only decrease, so where a plain read already shows lab[hi] <= lo the
// atomic would change nothing and is skipped: many entries join the same two
// fragments, and their atomics on one root's label would otherwise queue up
// behind each other (densMAP's 500k graph: 1.17 ms per sweep).
__device__ inline void cc_hook_pair(uint32_t u, uint32_t v, uint32_t* __restrict__ lab,
                                    uint32_t* __restrict__ changed) {
    uint32_t a = lab[u], b = lab[v];
    if (a == b) return;
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    if (__hip_atomic_load(&lab[hi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= lo) return;
    uint32_t old = atomicMin(&lab[hi], lo);
    if (old > lo) *changed = 1u;
}

// Sampling sweep: every vertex hooks to its first kCcSample neighbours.  On a
// kNN graph this joins most of each component at a tenth of a full sweep's
// gathers, so the full sweeps below mostly confirm.
constexpr int kCcSample = 2;
__global__ __launch_bounds__(kWg) void k_cc_sample(const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ col, uint32_t n,
                                                   uint32_t* __restrict__ lab,
                                                   uint32_t* __restrict__ changed) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = off[i], e = off[i + 1];
    for (uint32_t p = b; p < e && p < b + kCcSample; p++) cc_hook_pair(i, col[p], lab, changed);
}

// One hooking sweep over the filtered COO entries in row order.  (Sweeping
// only the undirected edge list, i < j, halves the gathers but took 2.2x as
// long at 1M nodes: in row order an entry (i, j < i) pulls j's label into
// row i, labels spread within a sweep, and most later entries see equal
// labels and skip their atomic.)
__global__ __launch_bounds__(kWg) void k_cc_hook(const uint32_t* __restrict__ row,
                                                 const uint32_t* __restrict__ col, uint64_t nnz,
                                                 uint32_t* __restrict__ lab,
                                                 uint32_t* __restrict__ changed) {
    uint64_t p = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (p >= nnz) return;
    cc_hook_pair(row[p], col[p], lab, changed);
}

__global__ __launch_bounds__(kWg) void k_cc_jump(uint32_t* __restrict__ lab, uint32_t n) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    uint32_t r = lab[i];
    while (lab[r] != r) r = lab[r];
    lab[i] = r;
}

// ---- component order on the device
// Components are numbered in ascending order of their smallest vertex, which
// is the root the labelling converges to; position q of the component order
// runs over the components one after another, each in ascending vertex order.

// flag[i] = 1 where i is a root (lab[i] == i).  *bad is raised where a label is
// out of range or does not point at a root (the labelling has not converged).
__global__ __launch_bounds__(kWg) void k_root_flags(const uint32_t* __restrict__ lab, uint32_t n,
                                                    uint32_t* __restrict__ flag,
                                                    uint32_t* __restrict__ bad) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = lab[i];
    if (r >= n || lab[r] != r) {
        *bad = 1u;
        flag[i] = 0u;
        return;
    }
    flag[i] = r == i ? 1u : 0u;
}

// comp[i] = the rank of i's root among the roots (rid = exclusive scan of the
// root flags).
__global__ __launch_bounds__(kWg) void k_comp_ids(const uint32_t* __restrict__ lab,
                                                  const uint32_t* __restrict__ rid, uint32_t n,
                                                  uint32_t* __restrict__ comp) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i < n) comp[i] = rid[lab[i]];
}

// start[c] = the first position of component c in the sorted ids; start[ncomp] = n
__global__ __launch_bounds__(kWg) void k_comp_starts(const uint32_t* __restrict__ skey, uint32_t n,
                                                     uint32_t* __restrict__ start) {
    const uint32_t q = blockIdx.x * kWg + threadIdx.x;
    if (q >= n) return;
    if (q == 0 || skey[q] != skey[q - 1]) start[skey[q]] = q;
    if (q == n - 1) start[skey[q] + 1] = n;
}

__global__ __launch_bounds__(kWg) void k_iota(uint32_t* __restrict__ a, uint32_t n) {
    const uint32_t q = blockIdx.x * kWg + threadIdx.x;
    if (q < n) a[q] = q;
}

// inv[perm[q]] = q
__global__ __launch_bounds__(kWg) void k_inverse(const uint32_t* __restrict__ perm, uint32_t n,
                                                 uint32_t* __restrict__ inv) {
    const uint32_t q = blockIdx.x * kWg + threadIdx.x;
    if (q < n) inv[perm[q]] = q;
}

// Degree of each vertex in component order, deg[q] = row length of perm[q],
// and its D^-1/2 in component order, dinv_p[q] = dinv[perm[q]].
__global__ __launch_bounds__(kWg) void k_perm_degree(const uint32_t* __restrict__ off,
                                                     const double* __restrict__ dinv,
                                                     const uint32_t* __restrict__ perm, uint32_t n,
                                                     uint32_t* __restrict__ deg,
                                                     double* __restrict__ dinv_p) {
    const uint32_t q = blockIdx.x * kWg + threadIdx.x;
    if (q >= n) return;
    const uint32_t i = perm[q];
    deg[q] = off[i + 1] - off[i];
    dinv_p[q] = dinv[i];
}

constexpr int kRowLanes = 16;  // lanes per CSR row in the row kernels below

// W in component order: row q is row perm[q] with its columns renamed through
// inv (= perm^-1), its weights kept in FP32 and the diagonal set to 0, entries
// in their row order.  Every Lanczos vector then lives in component order, so
// a component's vertices are one contiguous range.  16 lanes per row, each
// copying entries sub, sub + 16, ... of its row.
__global__ __launch_bounds__(kWg) void k_perm_rows(const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ col,
                                                   const float* __restrict__ val,
                                                   const uint32_t* __restrict__ perm,
                                                   const uint32_t* __restrict__ inv,
                                                   const uint32_t* __restrict__ poff, uint32_t n,
                                                   uint32_t* __restrict__ pcol,
                                                   float* __restrict__ pw) {
    const uint64_t t = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    const uint64_t q = t / kRowLanes;
    const uint32_t sub = (uint32_t)(t % kRowLanes);
    if (q >= n) return;
    const uint32_t i = perm[q];
    const uint32_t b = off[i], e = off[i + 1], o = poff[q];
    for (uint32_t p = b + sub; p < e; p += kRowLanes) {
        const uint32_t j = col[p];
        pcol[o + (p - b)] = inv[j];
        pw[o + (p - b)] = j == i ? 0.0f : val[p];
    }
}

// y = M x for M = -D^-1/2 W D^-1/2 (zero diagonal), from u = D^-1/2 x:
// y_q = -dinv_p[q] * sum_p w_p u[col_p].  The row streams 8 bytes per entry
// (FP32 weight, column) where precomputed FP64 products of M would take 12.
// kSpmvLanes lanes per row; lane s takes entries s, s + kSpmvLanes, ... in
// order, kSpmvUnroll of them per round with every column and weight load of
// the round issued together and then every gather (one dependent round trip
// per round instead of per entry: a kNN row of ~23 entries takes one round);
// the lanes' sums meet in a fixed xor tree, so the result does not depend on
// scheduling.  Workgroups go to the 8 XCDs round-robin, so block b works on
// row block (b mod 8) * (gridDim.x / 8) + b / 8: each XCD takes one
// contiguous eighth of the rows (in component order).  The grid is a
// multiple of 8 blocks (kSpmvXcds); blocks past the rows return.  Block 0
// also opens Lanczos step j (lanczos_gate): nothing in the product reads the
// liveness flags.
// Start of Lanczos step j for component q: a component past its iteration
// budget, or whose previous residual vanished (breakdown, math.rs:490-495),
// stops; a live one records that step j belongs to its basis.
__device__ inline void lanczos_gate(uint8_t* __restrict__ live, int* __restrict__ n_done,
                                    const int* __restrict__ n_iter,
                                    const double* __restrict__ beta_prev, uint32_t q, int j) {
    if (live[q] && (j >= n_iter[q] || (beta_prev && beta_prev[q] < 1e-12))) live[q] = 0;
    if (live[q]) n_done[q] = j + 1;
}

constexpr uint32_t kSpmvXcds = 8;
constexpr int kSpmvLanes = 8;
constexpr int kSpmvUnroll = 4;
__global__ __launch_bounds__(kWg) void k_spmv_rows(const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ col,
                                                   const float* __restrict__ w,
                                                   const double* __restrict__ u,
                                                   const double* __restrict__ dinv_p, uint32_t n,
                                                   double* __restrict__ y, uint8_t* __restrict__ live,
                                                   int* __restrict__ n_done,
                                                   const int* __restrict__ n_iter,
                                                   const double* __restrict__ beta_prev,
                                                   uint32_t nl, int j) {
    if (blockIdx.x == 0)
        for (uint32_t q = threadIdx.x; q < nl; q += kWg) lanczos_gate(live, n_done, n_iter, beta_prev, q, j);
    const uint32_t per_xcd = gridDim.x / kSpmvXcds;
    const uint32_t blk = (blockIdx.x % kSpmvXcds) * per_xcd + blockIdx.x / kSpmvXcds;
    const uint64_t t = (uint64_t)blk * kWg + threadIdx.x;
    const uint64_t row = t / kSpmvLanes;
    const uint32_t sub = (uint32_t)(t % kSpmvLanes);
    double acc = 0.0;
    if (row < n) {
        const uint32_t end = off[row + 1];
        for (uint32_t p0 = off[row] + sub; p0 < end; p0 += kSpmvLanes * kSpmvUnroll) {
            uint32_t c[kSpmvUnroll];
            float wv[kSpmvUnroll];
#pragma unroll
            for (int i = 0; i < kSpmvUnroll; i++) {
                const uint32_t p = p0 + (uint32_t)i * kSpmvLanes;
                c[i] = p < end ? col[p] : 0u;
                wv[i] = p < end ? w[p] : 0.0f;
            }
            double uv[kSpmvUnroll];
#pragma unroll
            for (int i = 0; i < kSpmvUnroll; i++)
                uv[i] = p0 + (uint32_t)i * kSpmvLanes < end ? u[c[i]] : 0.0;
#pragma unroll
            for (int i = 0; i < kSpmvUnroll; i++)
                if (p0 + (uint32_t)i * kSpmvLanes < end) acc += (double)wv[i] * uv[i];
        }
    }
#pragma unroll
    for (int o = kSpmvLanes / 2; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, kSpmvLanes);
    if (row < n && sub == 0) y[row] = -dinv_p[row] * acc;
}

struct Chunk {
    uint32_t comp, begin, end;
};

// Classical Gram-Schmidt, blocked: partial[chunk][k] = sum over the chunk's
// vertices of V_k[i] * w[i] for k in [k0, k0 + kKB), one block per (chunk, k-block).
constexpr int kKB = 8;
__global__ __launch_bounds__(kWg) void k_multidot(const Chunk* __restrict__ chunks,
                                                  const double* __restrict__ V, uint32_t n,
                                                  const double* __restrict__ w, int m,
                                                  double* __restrict__ partial) {
    __shared__ double sh[kKB][kWg];
    const Chunk c = chunks[blockIdx.x];
    const int k0 = blockIdx.y * kKB;
    double acc[kKB];
#pragma unroll
    for (int t = 0; t < kKB; t++) acc[t] = 0.0;
    for (uint32_t q = c.begin + threadIdx.x; q < c.end; q += kWg) {
        const double wq = w[q];
#pragma unroll
        for (int t = 0; t < kKB; t++)
            if (k0 + t < m) acc[t] += V[(size_t)(k0 + t) * n + q] * wq;
    }
#pragma unroll
    for (int t = 0; t < kKB; t++) sh[t][threadIdx.x] = acc[t];
    __syncthreads();
    for (int st = kWg / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
#pragma unroll
            for (int t = 0; t < kKB; t++) sh[t][threadIdx.x] += sh[t][threadIdx.x + st];
        }
        __syncthreads();
    }
    if ((int)threadIdx.x < kKB && k0 + (int)threadIdx.x < m)
        partial[(size_t)blockIdx.x * m + k0 + threadIdx.x] = sh[threadIdx.x][0];
}

// h[comp][k] = sum of the component's chunk partials, one wavefront per
// (comp, k): lane l adds chunks l, l+64, ... in order, then a fixed xor tree
// (the result does not depend on scheduling).  The projection on the newest
// basis vector (k = m-1) is the Lanczos diagonal coefficient: written to alpha
// (pass 0) or added to it (pass 1).
constexpr int kWavesPerWg = kWg / 64;
__global__ __launch_bounds__(kWg) void k_multidot_finish(const double* __restrict__ partial,
                                                         const uint32_t* __restrict__ comp_chunk0,
                                                         uint32_t n_comp, int m,
                                                         double* __restrict__ h,
                                                         double* __restrict__ alpha, int pass) {
    const uint32_t t = blockIdx.x * kWavesPerWg + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (t >= n_comp * (uint32_t)m) return;  // uniform per wavefront
    const uint32_t c = t / m, k = t % m;
    double s = 0.0;
    for (uint32_t q = comp_chunk0[c] + lane; q < comp_chunk0[c + 1]; q += 64)
        s += partial[(size_t)q * m + k];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) {
        h[t] = s;
        if ((int)k == m - 1) alpha[c] = pass ? alpha[c] + s : s;
    }
}

// w -= sum_k V_k h[comp][k] over one chunk of a live component (one block per
// chunk); with norm_partial, also the chunk's sum of the updated w^2.
__global__ __launch_bounds__(kWg) void k_cgs_update(const Chunk* __restrict__ chunks,
                                                    double* __restrict__ w,
                                                    const double* __restrict__ V, uint32_t n,
                                                    const double* __restrict__ h, int m,
                                                    const uint8_t* __restrict__ live,
                                                    double* __restrict__ norm_partial) {
    __shared__ double sh[kWg];
    const Chunk c = chunks[blockIdx.x];
    const bool on = live[c.comp] != 0;
    const double* hc = h + (size_t)c.comp * m;
    double acc = 0.0;
    for (uint32_t q = c.begin + threadIdx.x; q < c.end; q += kWg) {
        double x = w[q];
        if (on) {
            double s = 0.0;
            for (int k = 0; k < m; k++) s += V[(size_t)k * n + q] * hc[k];
            x -= s;
            w[q] = x;
        }
        acc += x * x;
    }
    if (!norm_partial) return;  // uniform per block
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = kWg / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) norm_partial[blockIdx.x] = sh[0];
}

// The coefficients h[k] (k < m) of one component from its chunks' partial
// dot products (chunks [c0, c1)), in the order of k_multidot_finish: lane l
// adds chunks c0 + l, c0 + l + 64, ... in order, then a fixed xor tree.  Wave
// w of the block takes k = w, w + kWavesPerWg, ...; the results go to sh_h.
__device__ inline void component_coefs(const double* __restrict__ partial, uint32_t c0,
                                       uint32_t c1, int m, double* sh_h) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = wave; k < m; k += kWavesPerWg) {
        double t = 0.0;
        for (uint32_t q = c0 + lane; q < c1; q += 64) t += partial[(size_t)q * m + k];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o);
        if (lane == 0) sh_h[k] = t;
    }
    __syncthreads();
}

// k_multidot_finish and k_cgs_update in one launch, for components of at most
// kFusedChunks chunks: each chunk's block sums its component's partials itself
// (the same fixed order, so the same bits) and the component's first chunk
// writes alpha.  With next_partial (the first pass), the block then computes
// the second pass's partial products of its chunk, V_k . w for k < m, from
// the rows it has just updated (each thread's rows stay in registers and its
// V entries come back from cache), in place of that pass's k_multidot: each
// thread adds its rows in order, a fixed xor tree joins a wave's lanes and the
// waves are added in order.  A small component's Lanczos step is then five
// launches instead of nine.
// (up to 131 072 rows: each block re-reads its component's partials, 128 x
// m doubles, which past that costs more than the launches it saves)
constexpr uint32_t kFusedChunks = 128;
constexpr int kMaxLanczos = 128;  // Lanczos steps: 4 * (n_comp + 1) + 30 <= 98
__global__ __launch_bounds__(kWg) void k_cgs_update_fused(const Chunk* __restrict__ chunks,
                                                          const uint32_t* __restrict__ comp_chunk0,
                                                          double* __restrict__ w,
                                                          const double* __restrict__ V, uint32_t n,
                                                          const double* __restrict__ mpart, int m,
                                                          const uint8_t* __restrict__ live,
                                                          double* __restrict__ alpha, int pass,
                                                          double* __restrict__ norm_partial,
                                                          double* __restrict__ next_partial) {
    constexpr int kRows = kChunk / kWg;  // rows per thread
    __shared__ double sh[kWg];
    __shared__ double sh_h[kMaxLanczos];
    __shared__ double sh_p[kWavesPerWg][kMaxLanczos];
    const Chunk c = chunks[blockIdx.x];
    const uint32_t c0 = comp_chunk0[c.comp];
    component_coefs(mpart, c0, comp_chunk0[c.comp + 1], m, sh_h);
    if (threadIdx.x == 0 && blockIdx.x == c0)
        alpha[c.comp] = pass ? alpha[c.comp] + sh_h[m - 1] : sh_h[m - 1];
    const bool on = live[c.comp] != 0;
    double acc = 0.0;
    double xs[kRows];
#pragma unroll
    for (int r = 0; r < kRows; r++) {
        const uint32_t q = c.begin + threadIdx.x + (uint32_t)r * kWg;
        xs[r] = 0.0;
        if (q < c.end) {
            double x = w[q];
            if (on) {
                double t = 0.0;
                for (int k = 0; k < m; k++) t += V[(size_t)k * n + q] * sh_h[k];
                x -= t;
                w[q] = x;
            }
            xs[r] = x;
            acc += x * x;
        }
    }
    if (next_partial) {  // uniform per block
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int k = 0; k < m; k++) {
            double t = 0.0;
#pragma unroll
            for (int r = 0; r < kRows; r++) {
                const uint32_t q = c.begin + threadIdx.x + (uint32_t)r * kWg;
                if (q < c.end) t += V[(size_t)k * n + q] * xs[r];
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o);
            if (lane == 0) sh_p[wave][k] = t;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < m; k += kWg) {
            double t = sh_p[0][k];
#pragma unroll
            for (int wv = 1; wv < kWavesPerWg; wv++) t += sh_p[wv][k];
            next_partial[(size_t)blockIdx.x * m + k] = t;
        }
    }
    if (!norm_partial) return;  // uniform per block
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = kWg / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) norm_partial[blockIdx.x] = sh[0];
}

// k_seg_finish (sqrt) and k_advance in one launch, one block per chunk, for
// components of at most kFusedChunks chunks: the block's first wave sums the
// component's norm partials (the order of k_seg_finish), the component's
// first chunk writes beta, and a live component's chunk rows advance.
__global__ __launch_bounds__(kWg) void k_advance_fused(const Chunk* __restrict__ chunks,
                                                       const uint32_t* __restrict__ comp_chunk0,
                                                       const double* __restrict__ partial,
                                                       double* __restrict__ beta,
                                                       double* __restrict__ u,
                                                       double* __restrict__ vnext,
                                                       const double* __restrict__ w,
                                                       const double* __restrict__ dinv_p,
                                                       const uint8_t* __restrict__ live) {
    __shared__ double sh_b;
    const Chunk c = chunks[blockIdx.x];
    const uint32_t c0 = comp_chunk0[c.comp];
    if (threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x;
        double t = 0.0;
        for (uint32_t q = c0 + lane; q < comp_chunk0[c.comp + 1]; q += 64) t += partial[q];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o);
        if (lane == 0) {
            sh_b = sqrt(t);
            if (blockIdx.x == c0) beta[c.comp] = sh_b;
        }
    }
    __syncthreads();
    if (!live[c.comp]) return;  // uniform per block
    const double b = sh_b;
    for (uint32_t i = c.begin + threadIdx.x; i < c.end; i += kWg) {
        const double x = w[i] / b;
        u[i] = x * dinv_p[i];
        if (vnext) vnext[i] = x;
    }
}

__device__ inline double block_sum(double v, double* sh) {
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

// partial[chunk] = sum over the chunk's vertices of a[i]*b[i]
__global__ __launch_bounds__(kWg) void k_seg_dot(const Chunk* __restrict__ chunks,
                                                 const double* __restrict__ a,
                                                 const double* __restrict__ b,
                                                 double* __restrict__ partial) {
    __shared__ double sh[kWg];
    const Chunk c = chunks[blockIdx.x];
    double acc = 0.0;
    for (uint32_t q = c.begin + threadIdx.x; q < c.end; q += kWg) acc += a[q] * b[q];
    double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[comp] = (sqrt?) sum of the comp's chunk partials, one wavefront per
// component, in the fixed order of k_multidot_finish
__global__ __launch_bounds__(kWg) void k_seg_finish(const double* __restrict__ partial,
                                                    const uint32_t* __restrict__ comp_chunk0,
                                                    uint32_t n_comp, int do_sqrt,
                                                    double* __restrict__ out) {
    const uint32_t c = blockIdx.x * kWavesPerWg + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (c >= n_comp) return;  // uniform per wavefront
    double t = 0.0;
    for (uint32_t q = comp_chunk0[c] + lane; q < comp_chunk0[c + 1]; q += 64) t += partial[q];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o);
    if (lane == 0) out[c] = do_sqrt ? sqrt(t) : t;
}

// v_{j+1} = w / beta_j into the next basis column V_{j+1}, and the next
// product's input u = D^-1/2 v_{j+1} (live components only)
__global__ __launch_bounds__(kWg) void k_advance(double* __restrict__ u, double* __restrict__ vnext,
                                                 const double* __restrict__ w,
                                                 const double* __restrict__ beta,
                                                 const double* __restrict__ dinv_p,
                                                 const uint32_t* __restrict__ comp_of,
                                                 const uint8_t* __restrict__ live, uint32_t n) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = comp_of[i];
    if (c == 0xFFFFFFFFu || !live[c]) return;
    const double x = w[i] / beta[c];
    u[i] = x * dinv_p[i];
    if (vnext) vnext[i] = x;
}

// u = D^-1/2 v (the first product's input; u may be v itself)
__global__ __launch_bounds__(kWg) void k_mul_dinv(const double* v,
                                                  const double* __restrict__ dinv_p, uint32_t n,
                                                  double* u) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i < n) u[i] = v[i] * dinv_p[i];
}


__global__ __launch_bounds__(kWg) void k_scale_seg(double* __restrict__ v,
                                                   const double* __restrict__ s,
                                                   const uint32_t* __restrict__ comp_of,
                                                   uint32_t n) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = comp_of[i];
    if (c == 0xFFFFFFFFu) return;
    v[i] = s[c] > 0.0 ? v[i] / s[c] : v[i];
}

// start vector in component order (vertex perm[q]'s draw at position q)
__global__ __launch_bounds__(kWg) void k_start_vec(double* __restrict__ v,
                                                   const uint32_t* __restrict__ perm, uint32_t n,
                                                   uint64_t seed) {
    uint32_t q = blockIdx.x * kWg + threadIdx.x;
    if (q < n) v[q] = u01(seed, 1, perm[q]) - 0.5;
}

// evec[i] = sum_j V_j[i] * C_comp(i)[j][sel]  for sel in [0, n_sel)
__global__ __launch_bounds__(kWg) void k_ritz(const double* __restrict__ V, uint32_t n,
                                              int max_iter, const double* __restrict__ coef,
                                              const uint32_t* __restrict__ coef_off,
                                              const int* __restrict__ n_done,
                                              const int* __restrict__ n_sel,
                                              const uint32_t* __restrict__ comp_of, int sel,
                                              double* __restrict__ out) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = comp_of[i];
    if (c == 0xFFFFFFFFu || sel >= n_sel[c]) return;
    const int m = n_done[c];
    const double* C = coef + coef_off[c];  // [m][n_sel]
    double acc = 0.0;
    for (int j = 0; j < m; j++) acc += V[(size_t)j * n + i] * C[j * n_sel[c] + sel];
    out[i] = acc;
}

// Position q's component by binary search of the component starts (start[c] <=
// q < start[c+1]; components are contiguous in component order).
__device__ inline uint32_t comp_at(const uint32_t* __restrict__ start, uint32_t ncomp, uint32_t q) {
    uint32_t lo = 0, hi = ncomp - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (start[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// lcq[q] = Lanczos index of position q's component (0xFFFFFFFF: not spectral)
__global__ __launch_bounds__(kWg) void k_fill_lcq(const uint32_t* __restrict__ start,
                                                  uint32_t ncomp,
                                                  const uint32_t* __restrict__ comp_to_lc,
                                                  uint32_t n, uint32_t* __restrict__ lcq) {
    const uint32_t q = blockIdx.x * kWg + threadIdx.x;
    if (q < n) lcq[q] = comp_to_lc[comp_at(start, ncomp, q)];
}

// raw[q][dcomp] = the normalised Ritz vector (f32, eigenvectors.push(T::from),
// math.rs:583), or a uniform(-1, 1) fill where the component has fewer
// converged pairs than requested.
__global__ __launch_bounds__(kWg) void k_raw_store(const double* __restrict__ evec,
                                                   const uint32_t* __restrict__ lcq,
                                                   const uint32_t* __restrict__ perm,
                                                   const int* __restrict__ n_sel, int sel,
                                                   int dcomp, int n_comp, uint64_t seed,
                                                   uint32_t n, float* __restrict__ raw) {
    const uint32_t q = blockIdx.x * kWg + threadIdx.x;
    if (q >= n) return;
    const uint32_t c = lcq[q];
    if (c == 0xFFFFFFFFu) return;
    raw[(size_t)q * n_comp + dcomp] =
        sel < n_sel[c] ? (float)evec[q] : (float)(2.0 * u01(seed + dcomp, 3, perm[q]) - 1.0);
}

// partial[chunk] = max |raw| over the chunk's rows (all n_comp coordinates)
__global__ __launch_bounds__(kWg) void k_chunk_absmax(const Chunk* __restrict__ chunks,
                                                      const float* __restrict__ raw, int n_comp,
                                                      float* __restrict__ partial) {
    __shared__ float sh[kWg];
    const Chunk c = chunks[blockIdx.x];
    float m = 0.0f;
    for (uint32_t q = c.begin + threadIdx.x; q < c.end; q += kWg)
        for (int dd = 0; dd < n_comp; dd++) m = fmaxf(m, fabsf(raw[(size_t)q * n_comp + dd]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int st = kWg / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// multi_component_init placement (init.rs:439-484): a spectral component's
// rows are meta + raw * expansion, the others meta + uniform(-1, 1) * range.
__global__ __launch_bounds__(kWg) void k_place(const float* __restrict__ raw,
                                               const uint32_t* __restrict__ start, uint32_t ncomp,
                                               const uint32_t* __restrict__ perm,
                                               const float* __restrict__ meta,
                                               const float* __restrict__ expansion,
                                               const float* __restrict__ drange,
                                               const uint8_t* __restrict__ spectral, uint32_t n,
                                               int n_comp, uint64_t seed,
                                               float* __restrict__ emb) {
    const uint32_t q = blockIdx.x * kWg + threadIdx.x;
    if (q >= n) return;
    const uint32_t c = comp_at(start, ncomp, q);
    const uint32_t i = perm[q];
    for (int dd = 0; dd < n_comp; dd++) {
        const size_t o = (size_t)q * n_comp + dd;
        const float mc = meta[(size_t)c * n_comp + dd];
        if (spectral[c]) {
            emb[o] = mc + raw[o] * expansion[c];
        } else {
            const float u = (float)(2.0 * u01(seed, 19, (uint64_t)i * n_comp + dd) - 1.0);
            emb[o] = mc + u * drange[c];
        }
    }
}

// a single component too small for the spectral branch: uniform(-1, 1)
__global__ __launch_bounds__(kWg) void k_random_rows(uint32_t n, int n_comp, uint64_t seed,
                                                     float* __restrict__ emb) {
    const uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    for (int dd = 0; dd < n_comp; dd++)
        emb[(size_t)i * n_comp + dd] = (float)(2.0 * u01(seed, 5, (uint64_t)i * n_comp + dd) - 1.0);
}

constexpr int kFinBlocks = 256;  // fixed grid of the finishing reductions (run-to-run identical)

// per-block FP64 column sums of emb [n][n_comp]
__global__ __launch_bounds__(kWg) void k_col_sums(const float* __restrict__ emb, uint32_t n,
                                                  int n_comp, double* __restrict__ partials) {
    __shared__ double sh[kWg];
    for (int dd = 0; dd < n_comp; dd++) {
        double acc = 0.0;
        for (uint32_t q = blockIdx.x * kWg + threadIdx.x; q < n; q += kFinBlocks * kWg)
            acc += (double)emb[(size_t)q * n_comp + dd];
        const double t = block_sum(acc, sh);
        if (threadIdx.x == 0) partials[(size_t)blockIdx.x * n_comp + dd] = t;
    }
}

// centre every column on its mean (f32, the mean rounded from the FP64 sum)
// and emit per-block max |x|
__global__ __launch_bounds__(kWg) void k_center_absmax(float* __restrict__ emb, uint32_t n,
                                                       int n_comp,
                                                       const double* __restrict__ partials,
                                                       float* __restrict__ absmax) {
    __shared__ float mean[16];
    __shared__ float sh[kWg];
    if ((int)threadIdx.x < n_comp) {
        double t = 0.0;
        for (int b = 0; b < kFinBlocks; b++) t += partials[(size_t)b * n_comp + threadIdx.x];
        mean[threadIdx.x] = (float)(t / n);
    }
    __syncthreads();
    float m = 0.0f;
    for (uint32_t q = blockIdx.x * kWg + threadIdx.x; q < n; q += kFinBlocks * kWg)
        for (int dd = 0; dd < n_comp; dd++) {
            const size_t o = (size_t)q * n_comp + dd;
            const float x = emb[o] - mean[dd];
            emb[o] = x;
            m = fmaxf(m, fabsf(x));
        }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int st = kWg / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) absmax[blockIdx.x] = sh[0];
}

// scale max |x| to `range`, optionally add the N(0, 1e-4) noise of
// finalise_spectral_embedding (Box-Muller on counter hashes), and write row q
// to the caller's row perm[q]
__global__ __launch_bounds__(kWg) void k_finish_rows(const float* __restrict__ emb, uint32_t n,
                                                     int n_comp,
                                                     const float* __restrict__ absmax,
                                                     float range, int noise, uint64_t seed,
                                                     const uint32_t* __restrict__ perm,
                                                     float* __restrict__ out) {
    __shared__ float mx_sh;
    if (threadIdx.x == 0) {
        float mx = 0.0f;
        for (int b = 0; b < kFinBlocks; b++) mx = fmaxf(mx, absmax[b]);
        mx_sh = mx;
    }
    __syncthreads();
    const float mx = mx_sh;
    const bool scale = mx > 1e-8f;
    const float sc = scale ? range / mx : 1.0f;
    const uint32_t q = blockIdx.x * kWg + threadIdx.x;
    if (q >= n) return;
    const uint32_t i = perm[q];
    for (int dd = 0; dd < n_comp; dd++) {
        float x = emb[(size_t)q * n_comp + dd];
        if (scale) x *= sc;
        if (noise) {
            const uint64_t f = (uint64_t)i * n_comp + dd;
            const double u1 = fmax(u01(seed + 9999, 11, f), 1e-300), u2 = u01(seed + 9999, 13, f);
            const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
            x += (float)(z * 1e-4);
        }
        out[(size_t)i * n_comp + dd] = x;
    }
}

// ---------------------------------------------------------- host helpers

}  // namespace

// Eigen-decomposition of the symmetric tridiagonal matrix with diagonal
// alpha[0..m) and off-diagonal beta[0..m-1): implicit QL iterations with
// Wilkinson shifts, the rotations accumulated into the eigenvector matrix
// (Golub & Van Loan, Matrix Computations, 4th ed., section 8.3.3: each sweep
// chases the bulge of a shifted QL step from the bottom of the unreduced block
// up to its top with Givens rotations).  Returns eigenvalues ascending and the
// eigenvectors as columns of vecs [m][m] (row-major, vecs[k*m + r] =
// component k of eigenvector r).  O(m^2) per eigenvalue.
void tridiag_eig(const std::vector<double>& alpha, const std::vector<double>& beta, int m,
                 std::vector<double>& evals, std::vector<double>& vecs) {
    std::vector<double> d(alpha.begin(), alpha.begin() + m), e((size_t)m, 0.0);
    for (int i = 0; i + 1 < m; i++) e[i] = beta[i];
    std::vector<double> Z((size_t)m * m, 0.0);
    for (int i = 0; i < m; i++) Z[(size_t)i * m + i] = 1.0;
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < m; l++) {
        for (int iter = 0; iter < 64; iter++) {
            // end of the unreduced block that starts at l
            int hi = l;
            while (hi + 1 < m &&
                   std::fabs(e[hi]) > eps * (std::fabs(d[hi]) + std::fabs(d[hi + 1])))
                hi++;
            if (hi == l) break;  // d[l] has converged
            // Wilkinson shift from the leading 2x2 block
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[hi] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double sn = 1.0, cs = 1.0, p = 0.0;
            bool deflated = false;
            for (int i = hi - 1; i >= l; i--) {
                const double f = sn * e[i], bb = cs * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {  // an exact split: the block below i is done
                    d[i + 1] -= p;
                    e[hi] = 0.0;
                    deflated = true;
                    break;
                }
                sn = f / r;
                cs = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * sn + 2.0 * cs * bb;
                p = sn * r;
                d[i + 1] = g + p;
                g = cs * r - bb;
                for (int k = 0; k < m; k++) {
                    const double zi = Z[(size_t)k * m + i], zi1 = Z[(size_t)k * m + i + 1];
                    Z[(size_t)k * m + i + 1] = sn * zi + cs * zi1;
                    Z[(size_t)k * m + i] = cs * zi - sn * zi1;
                }
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[hi] = 0.0;
        }
    }
    std::vector<int> order(m);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
    evals.resize(m);
    vecs.assign((size_t)m * m, 0.0);
    for (int r2 = 0; r2 < m; r2++) {
        evals[r2] = d[order[r2]];
        for (int k = 0; k < m; k++) vecs[(size_t)k * m + r2] = Z[(size_t)k * m + order[r2]];
    }
}

namespace {

// Host Lanczos for a small dense symmetric graph (the component meta graph):
// smallest n_eig eigenvectors of -D^-1/2 W D^-1/2, as [n][n_sel] (f32 values).
std::vector<std::vector<float>> host_small_spectral(const std::vector<double>& W, int n,
                                                    int n_comp, uint64_t seed) {
    std::vector<std::vector<float>> emb(n, std::vector<float>(n_comp, 0.0f));
    if (n <= n_comp + 1) {
        for (int i = 0; i < n; i++)
            for (int d = 0; d < n_comp; d++) emb[i][d] = (float)(2.0 * u01(seed, 7, i * 16 + d) - 1.0);
        return emb;
    }
    std::vector<double> dinv(n);
    for (int i = 0; i < n; i++) {
        double s = 0;
        for (int j = 0; j < n; j++) s += W[(size_t)i * n + j];
        dinv[i] = s > 1e-8 ? 1.0 / std::sqrt(s) : 0.0;
    }
    auto mv = [&](const std::vector<double>& x, std::vector<double>& y) {
        for (int i = 0; i < n; i++) {
            double a = 0;
            for (int j = 0; j < n; j++)
                if (j != i) a += (-dinv[i] * W[(size_t)i * n + j] * dinv[j]) * x[j];
            y[i] = a;
        }
    };
    const int n_eig = std::min(n_comp + 1, n);
    const int n_iter = std::min(n_eig * 4 + 30, n);
    std::vector<double> v(n), v_old(n, 0.0), w(n), basis;
    for (int i = 0; i < n; i++) v[i] = u01(seed, 1, i) - 0.5;
    double nv = 0;
    for (double x : v) nv += x * x;
    nv = std::sqrt(nv);
    for (double& x : v) x /= nv;
    std::vector<double> alpha, beta;
    int done = 0;
    for (int j = 0; j < n_iter; j++) {
        basis.insert(basis.end(), v.begin(), v.end());
        done = j + 1;
        mv(v, w);
        double a = 0;
        for (int i = 0; i < n; i++) a += w[i] * v[i];
        alpha.push_back(a);
        for (int i = 0; i < n; i++) {
            w[i] -= a * v[i];
            if (j > 0) w[i] -= beta[j - 1] * v_old[i];
        }
        for (int k = 0; k <= j; k++) {
            double c = 0;
            for (int i = 0; i < n; i++) c += w[i] * basis[(size_t)k * n + i];
            for (int i = 0; i < n; i++) w[i] -= c * basis[(size_t)k * n + i];
        }
        double b = 0;
        for (double x : w) b += x * x;
        b = std::sqrt(b);
        beta.push_back(b);
        if (b < 1e-12) break;
        v_old = v;
        for (int i = 0; i < n; i++) v[i] = w[i] / b;
    }
    std::vector<double> ev, vecs;
    tridiag_eig(alpha, beta, done, ev, vecs);
    const int n_sel = std::min(n_eig, done);
    for (int d = 0; d < n_comp; d++) {
        const int s = d + 1;
        if (s < n_sel) {
            std::vector<double> e(n, 0.0);
            for (int j = 0; j < done; j++)
                for (int i = 0; i < n; i++) e[i] += basis[(size_t)j * n + i] * vecs[j * done + s];
            double nn = 0;
            for (double x : e) nn += x * x;
            nn = std::sqrt(nn);
            for (int i = 0; i < n; i++) emb[i][d] = (float)(nn > 0 ? e[i] / nn : e[i]);
        } else {
            for (int i = 0; i < n; i++) emb[i][d] = (float)(2.0 * u01(seed + d, 3, i) - 1.0);
        }
    }
    return emb;
}

__global__ __launch_bounds__(kWg) void k_centroid_partials(const Chunk* __restrict__ chunks,
                                                           const uint32_t* __restrict__ perm,
                                                           const float* __restrict__ x, int d,
                                                           double* __restrict__ partial) {
    const Chunk c = chunks[blockIdx.x];
    for (int f = threadIdx.x; f < d; f += kWg) {
        double acc = 0.0;
        for (uint32_t q = c.begin; q < c.end; q++) acc += (double)x[(size_t)perm[q] * d + f];
        partial[(size_t)blockIdx.x * d + f] = acc;
    }
}

}  // namespace

void spectral_layout_dev(mrs_ctx* ctx, const mrs_graph* g, int32_t dim, uint64_t seed,
                         float range, float* out, const float* data, int32_t d) {
    const uint32_t n = (uint32_t)g->n;
    const uint64_t nnz = (uint64_t)g->nnz;
    hipStream_t s = ctx->stream;
    const int n_comp = dim;

    // ---- CSR view of the filtered graph and D^-1/2
    DevBuf<uint32_t> off((size_t)n + 1);
    hipLaunchKernelGGL(k_row_offsets, dim3(div_up((uint64_t)n + 1, kWg)), dim3(kWg), 0, s,
                       g->coo_row.p, nnz, n, off.p);
    DevBuf<double> dinv(n);
    hipLaunchKernelGGL(k_degrees, dim3(div_up(n, kWg)), dim3(kWg), 0, s, off.p, g->coo_val.p, n,
                       dinv.p);
    MRS_CHECK_LAUNCH();

    // ---- connected components
    DevBuf<uint32_t> lab(n), changed(1);
    hipLaunchKernelGGL(k_cc_init, dim3(div_up(n, kWg)), dim3(kWg), 0, s, lab.p, n);
    if (nnz) {
        hipLaunchKernelGGL(k_cc_sample, dim3(div_up(n, kWg)), dim3(kWg), 0, s, off.p,
                           g->coo_col.p, n, lab.p, changed.p);
        hipLaunchKernelGGL(k_cc_jump, dim3(div_up(n, kWg)), dim3(kWg), 0, s, lab.p, n);
        MRS_CHECK_LAUNCH();
    }
    for (int it = 0; it < 10000; it++) {
        MRS_HIP(hipMemsetAsync(changed.p, 0, sizeof(uint32_t), s));
        if (nnz)
            hipLaunchKernelGGL(k_cc_hook, dim3(div_up(nnz, kWg)), dim3(kWg), 0, s, g->coo_row.p,
                               g->coo_col.p, nnz, lab.p, changed.p);
        hipLaunchKernelGGL(k_cc_jump, dim3(div_up(n, kWg)), dim3(kWg), 0, s, lab.p, n);
        MRS_CHECK_LAUNCH();
        uint32_t ch = 0;
        MRS_HIP(hipMemcpyAsync(&ch, changed.p, sizeof(ch), hipMemcpyDeviceToHost, s));
        MRS_HIP(hipStreamSynchronize(s));
        if (!ch) break;
    }
    // Component order, built on the device (a stable sort of the component
    // ids carrying the vertex ids): perm[q] is
    // the vertex at position q and inv its inverse, so a component is the
    // contiguous range [start[c], start[c+1]) and every vector access is a
    // contiguous stream.
    DevBuf<uint32_t> rid((size_t)n + 1), bad(1);
    MRS_HIP(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_root_flags, dim3(div_up(n, kWg)), dim3(kWg), 0, s, lab.p, n, rid.p, bad.p);
    MRS_CHECK_LAUNCH();
    exclusive_scan_u32(rid.p, rid.p, n, rid.p + n, s);
    uint32_t hdr[2] = {0, 0};
    MRS_HIP(hipMemcpyAsync(&hdr[0], rid.p + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MRS_HIP(hipMemcpyAsync(&hdr[1], bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MRS_HIP(hipStreamSynchronize(s));
    const uint32_t ncomp = hdr[0];
    if (hdr[1] || ncomp == 0 || ncomp > n)
        throw Error(MRS_DEVICE_ERROR, "connected components did not converge");
    DevBuf<uint32_t> dperm(n), dinvp(n), dstart((size_t)ncomp + 1);
    std::vector<uint32_t> start(ncomp + 1, 0);
    if (ncomp == 1) {
        hipLaunchKernelGGL(k_iota, dim3(div_up(n, kWg)), dim3(kWg), 0, s, dperm.p, n);
        hipLaunchKernelGGL(k_iota, dim3(div_up(n, kWg)), dim3(kWg), 0, s, dinvp.p, n);
        start[1] = n;
        MRS_HIP(hipMemcpyAsync(dstart.p, start.data(), sizeof(uint32_t) * 2, hipMemcpyHostToDevice, s));
        MRS_CHECK_LAUNCH();
    } else {
        DevBuf<uint32_t> ka(n), kb(n);
        hipLaunchKernelGGL(k_comp_ids, dim3(div_up(n, kWg)), dim3(kWg), 0, s, lab.p, rid.p, n, ka.p);
        MRS_CHECK_LAUNCH();
        radix_sort_pairs_u32(ka.p, nullptr, n, 32 - __builtin_clz(ncomp - 1), kb.p, dperm.p, s);
        const uint32_t* kin = kb.p;
        hipLaunchKernelGGL(k_comp_starts, dim3(div_up(n, kWg)), dim3(kWg), 0, s, kin, n, dstart.p);
        hipLaunchKernelGGL(k_inverse, dim3(div_up(n, kWg)), dim3(kWg), 0, s, dperm.p, n, dinvp.p);
        MRS_CHECK_LAUNCH();
        MRS_HIP(hipMemcpyAsync(start.data(), dstart.p, sizeof(uint32_t) * (ncomp + 1),
                               hipMemcpyDeviceToHost, s));
        MRS_HIP(hipStreamSynchronize(s));
    }
    std::vector<uint32_t> size(ncomp);
    for (uint32_t c = 0; c < ncomp; c++) {
        if (start[c + 1] <= start[c]) throw Error(MRS_DEVICE_ERROR, "empty connected component");
        size[c] = start[c + 1] - start[c];
    }

    // which components take the spectral branch
    const bool single = ncomp == 1;
    std::vector<uint8_t> spectral(ncomp, 0);
    for (uint32_t c = 0; c < ncomp; c++)
        spectral[c] = single ? (size[c] > (uint32_t)n_comp + 1)
                             : (size[c] >= 2u * n_comp && size[c] > (uint32_t)n_comp + 1);
    std::vector<uint32_t> lc_comp;  // lanczos comp -> component
    for (uint32_t c = 0; c < ncomp; c++)
        if (spectral[c]) lc_comp.push_back(c);
    const uint32_t nl = (uint32_t)lc_comp.size();
    std::vector<uint32_t> comp_to_lc(ncomp, 0xFFFFFFFFu);
    for (uint32_t q = 0; q < nl; q++) comp_to_lc[lc_comp[q]] = q;
    // reduction chunks of the spectral components (fixed: results are run-to-run identical)
    std::vector<Chunk> chunks;
    std::vector<uint32_t> chunk0(nl + 1, 0);
    for (uint32_t q = 0; q < nl; q++) {
        const uint32_t c = lc_comp[q];
        chunk0[q] = (uint32_t)chunks.size();
        for (uint32_t b = start[c]; b < start[c + 1]; b += kChunk)
            chunks.push_back({q, b, std::min(b + kChunk, start[c + 1])});
    }
    chunk0[nl] = (uint32_t)chunks.size();
    const uint32_t nch = (uint32_t)chunks.size();

    // raw spectral coordinates [n][n_comp] (f32), component order
    DevBuf<float> raw((size_t)n * n_comp);
    MRS_HIP(hipMemsetAsync(raw.p, 0, raw.bytes(), s));
    DevBuf<Chunk> dchunks(nch ? nch : 1);
    if (nl) {
        const int n_eig = n_comp + 1;
        std::vector<int> n_iter_c(nl), n_done(nl, 0), n_sel(nl, 0);
        int max_iter = 0;
        for (uint32_t q = 0; q < nl; q++) {
            n_iter_c[q] = (int)std::min<uint32_t>(n_eig * 4 + 30, size[lc_comp[q]]);
            max_iter = std::max(max_iter, n_iter_c[q]);
        }
        DevBuf<uint32_t> dchunk0(nl + 1), dc2l(ncomp), dlc(n), poff((size_t)n + 1);
        MRS_HIP(hipMemcpyAsync(dchunks.p, chunks.data(), sizeof(Chunk) * nch, hipMemcpyHostToDevice, s));
        MRS_HIP(hipMemcpyAsync(dchunk0.p, chunk0.data(), sizeof(uint32_t) * (nl + 1), hipMemcpyHostToDevice, s));
        MRS_HIP(hipMemcpyAsync(dc2l.p, comp_to_lc.data(), sizeof(uint32_t) * ncomp, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_fill_lcq, dim3(div_up(n, kWg)), dim3(kWg), 0, s, dstart.p, ncomp,
                           dc2l.p, n, dlc.p);
        // W and D^-1/2 in component order for M = -D^-1/2 W D^-1/2 (rows keep
        // their entry order, so every row sum adds its terms in a fixed order)
        DevBuf<double> dinv_p(n);
        hipLaunchKernelGGL(k_perm_degree, dim3(div_up(n, kWg)), dim3(kWg), 0, s, off.p, dinv.p,
                           dperm.p, n, poff.p, dinv_p.p);
        MRS_CHECK_LAUNCH();
        exclusive_scan_u32(poff.p, poff.p, n, poff.p + n, s);
        DevBuf<uint32_t> pcol(nnz ? nnz : 1);
        DevBuf<float> pw(nnz ? nnz : 1);
        hipLaunchKernelGGL(k_perm_rows, dim3(div_up((uint64_t)n * kRowLanes, kWg)), dim3(kWg), 0,
                           s, off.p, g->coo_col.p, g->coo_val.p, dperm.p, dinvp.p, poff.p, n,
                           pcol.p, pw.p);
        MRS_CHECK_LAUNCH();
        DevBuf<double> v(n), w(n), V((size_t)max_iter * n), partial(nch), sc(nl),
            alpha_d((size_t)max_iter * nl), beta_d((size_t)max_iter * nl),
            mpart((size_t)nch * max_iter), mpart2((size_t)nch * max_iter),
            hcoef((size_t)nl * max_iter);
        // per-component control stays on the device: the loop below issues no
        // host synchronisation until the tridiagonal coefficients are read back
        DevBuf<uint8_t> live(nl);
        DevBuf<int> dn_iter(nl), dn_done(nl);
        MRS_HIP(hipMemsetAsync(live.p, 1, nl, s));
        MRS_HIP(hipMemsetAsync(dn_done.p, 0, dn_done.bytes(), s));
        MRS_HIP(hipMemcpyAsync(dn_iter.p, n_iter_c.data(), sizeof(int) * nl, hipMemcpyHostToDevice, s));
        const dim3 gn(div_up(n, kWg)), gqw(div_up(nl, kWavesPerWg));
        auto seg = [&](const double* a, const double* b, double* outp, int do_sqrt) {
            hipLaunchKernelGGL(k_seg_dot, dim3(nch), dim3(kWg), 0, s, dchunks.p, a, b, partial.p);
            hipLaunchKernelGGL(k_seg_finish, gqw, dim3(kWg), 0, s, partial.p, dchunk0.p, nl,
                               do_sqrt, outp);
        };
        hipLaunchKernelGGL(k_start_vec, gn, dim3(kWg), 0, s, v.p, dperm.p, n, seed);
        seg(v.p, v.p, sc.p, 1);
        hipLaunchKernelGGL(k_scale_seg, gn, dim3(kWg), 0, s, v.p, sc.p, dlc.p, n);
        MRS_HIP(hipMemcpyAsync(V.p, v.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_mul_dinv, gn, dim3(kWg), 0, s, v.p, dinv_p.p, n, v.p);
        MRS_CHECK_LAUNCH();
        const unsigned spmv_blocks =
            div_up(div_up((uint64_t)n * kSpmvLanes, kWg), kSpmvXcds) * kSpmvXcds;
        // components of at most kFusedChunks chunks take the fused kernels
        // (five launches per step instead of nine)
        uint32_t max_chunks = 0;
        for (uint32_t q = 0; q < nl; q++) max_chunks = std::max(max_chunks, chunk0[q + 1] - chunk0[q]);
        const bool fused = max_chunks <= kFusedChunks;
        if (max_iter > kMaxLanczos) throw Error(MRS_INVALID_ARGUMENT, "too many Lanczos steps");
        for (int j = 0; j < max_iter; j++) {
            double* alpha_j = alpha_d.p + (size_t)j * nl;
            double* beta_j = beta_d.p + (size_t)j * nl;
            hipLaunchKernelGGL(k_spmv_rows, dim3(spmv_blocks), dim3(kWg), 0, s, poff.p, pcol.p,
                               pw.p, v.p, dinv_p.p, n, w.p, live.p, dn_done.p, dn_iter.p,
                               j > 0 ? beta_j - nl : nullptr, nl, j);
            // Orthogonalise w = M v_j against V_0..V_j by classical Gram-Schmidt
            // applied twice.  The first pass's projections on v_j and v_{j-1}
            // are alpha_j and beta_{j-1}, so it also performs the three-term
            // recurrence (math.rs:478-483); alpha_j collects both passes' v_j
            // projections, and the second pass emits the norm partials of w.
            const int m = j + 1;
            for (int pass = 0; pass < 2; pass++) {
                // fused: the second pass's partial products come from the first
                // pass's update (into mpart2)
                if (!fused || pass == 0)
                    hipLaunchKernelGGL(k_multidot, dim3(nch, div_up(m, kKB)), dim3(kWg), 0, s,
                                       dchunks.p, V.p, n, w.p, m, mpart.p);
                if (fused) {
                    hipLaunchKernelGGL(k_cgs_update_fused, dim3(nch), dim3(kWg), 0, s, dchunks.p,
                                       dchunk0.p, w.p, V.p, n, pass ? mpart2.p : mpart.p, m,
                                       live.p, alpha_j, pass, pass ? partial.p : nullptr,
                                       pass ? nullptr : mpart2.p);
                } else {
                    hipLaunchKernelGGL(k_multidot_finish,
                                       dim3(div_up((uint64_t)nl * m, kWavesPerWg)), dim3(kWg), 0,
                                       s, mpart.p, dchunk0.p, nl, m, hcoef.p, alpha_j, pass);
                    hipLaunchKernelGGL(k_cgs_update, dim3(nch), dim3(kWg), 0, s, dchunks.p, w.p,
                                       V.p, n, hcoef.p, m, live.p, pass ? partial.p : nullptr);
                }
            }
            double* vnext = j + 1 < max_iter ? V.p + (size_t)(j + 1) * n : nullptr;
            if (fused) {
                hipLaunchKernelGGL(k_advance_fused, dim3(nch), dim3(kWg), 0, s, dchunks.p,
                                   dchunk0.p, partial.p, beta_j, v.p, vnext, w.p, dinv_p.p, live.p);
            } else {
                hipLaunchKernelGGL(k_seg_finish, gqw, dim3(kWg), 0, s, partial.p, dchunk0.p, nl, 1,
                                   beta_j);
                hipLaunchKernelGGL(k_advance, gn, dim3(kWg), 0, s, v.p, vnext, w.p, beta_j,
                                   dinv_p.p, dlc.p, live.p, n);
            }
            MRS_CHECK_LAUNCH();
        }
        MRS_HIP(hipMemcpyAsync(n_done.data(), dn_done.p, sizeof(int) * nl, hipMemcpyDeviceToHost, s));
        std::vector<double> ha((size_t)max_iter * nl), hb((size_t)max_iter * nl);
        MRS_HIP(hipMemcpyAsync(ha.data(), alpha_d.p, sizeof(double) * ha.size(), hipMemcpyDeviceToHost, s));
        MRS_HIP(hipMemcpyAsync(hb.data(), beta_d.p, sizeof(double) * hb.size(), hipMemcpyDeviceToHost, s));
        MRS_HIP(hipStreamSynchronize(s));
        // tridiagonal problems on the host, one per component and independent
        // of each other: spread over a few threads (a 1M-point fit with 50
        // components spent 2.5 ms here on one thread); Ritz coefficients for
        // the smallest pairs are then gathered in component order
        std::vector<std::vector<double>> vecs_q(nl);
        {
            const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
            const unsigned nthr = nl >= 8 ? std::min(std::min(hw, 8u), nl) : 1u;
            std::vector<std::exception_ptr> errs(nthr);
            auto work = [&](unsigned t) {
                try {
                    for (uint32_t q = t; q < nl; q += nthr) {
                        const int m = n_done[q];
                        std::vector<double> al(m), be(std::max(m - 1, 0)), ev;
                        for (int j = 0; j < m; j++) al[j] = ha[(size_t)j * nl + q];
                        for (int j = 0; j + 1 < m; j++) be[j] = hb[(size_t)j * nl + q];
                        tridiag_eig(al, be, m, ev, vecs_q[q]);
                    }
                } catch (...) {
                    errs[t] = std::current_exception();
                }
            };
            std::vector<std::thread> pool;
            for (unsigned t = 1; t < nthr; t++) pool.emplace_back(work, t);
            work(0);
            for (auto& th : pool) th.join();
            for (auto& e : errs)
                if (e) std::rethrow_exception(e);
        }
        std::vector<double> coef;
        std::vector<uint32_t> coef_off(nl, 0);
        for (uint32_t q = 0; q < nl; q++) {
            const int m = n_done[q];
            const std::vector<double>& vecs = vecs_q[q];
            n_sel[q] = std::min(n_eig, m);
            coef_off[q] = (uint32_t)coef.size();
            for (int j = 0; j < m; j++)
                for (int t = 0; t < n_sel[q]; t++) coef.push_back(vecs[j * m + t]);
        }
        DevBuf<double> dcoef(coef.size() ? coef.size() : 1), evec(n);
        DevBuf<uint32_t> dcoef_off(nl);
        DevBuf<int> dn_sel(nl);
        MRS_HIP(hipMemcpyAsync(dcoef.p, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, s));
        MRS_HIP(hipMemcpyAsync(dcoef_off.p, coef_off.data(), sizeof(uint32_t) * nl, hipMemcpyHostToDevice, s));
        MRS_HIP(hipMemcpyAsync(dn_sel.p, n_sel.data(), sizeof(int) * nl, hipMemcpyHostToDevice, s));
        for (int dcomp = 0; dcomp < n_comp; dcomp++) {
            const int sel = dcomp + 1;  // skip the trivial smallest eigenvector
            MRS_HIP(hipMemsetAsync(evec.p, 0, evec.bytes(), s));
            hipLaunchKernelGGL(k_ritz, gn, dim3(kWg), 0, s, V.p, n, max_iter, dcoef.p,
                               dcoef_off.p, dn_done.p, dn_sel.p, dlc.p, sel, evec.p);
            seg(evec.p, evec.p, sc.p, 1);
            hipLaunchKernelGGL(k_scale_seg, gn, dim3(kWg), 0, s, evec.p, sc.p, dlc.p, n);
            hipLaunchKernelGGL(k_raw_store, gn, dim3(kWg), 0, s, evec.p, dlc.p, dperm.p, dn_sel.p,
                               sel, dcomp, n_comp, seed, n, raw.p);
            MRS_CHECK_LAUNCH();
        }
        // the stream is synchronised before these buffers are freed
        MRS_HIP(hipStreamSynchronize(s));
    }

    DevBuf<float> emb((size_t)n * n_comp);
    const dim3 gn(div_up(n, kWg));
    int noise = 0;
    if (single) {
        // finalise_spectral_embedding (init.rs:905-960)
        if (!spectral[0])
            hipLaunchKernelGGL(k_random_rows, gn, dim3(kWg), 0, s, n, n_comp, seed, emb.p);
        else
            MRS_HIP(hipMemcpyAsync(emb.p, raw.p, raw.bytes(), hipMemcpyDeviceToDevice, s));
        noise = 1;
    } else {
        // component_meta_layout (:539-595)
        std::vector<std::vector<float>> meta(ncomp, std::vector<float>(n_comp, 0.0f));
        auto simplex = [&]() {
            const uint32_t kk = (ncomp + 1) / 2;
            for (auto& r : meta) std::fill(r.begin(), r.end(), 0.0f);
            for (uint32_t l = 0; l < ncomp; l++) {
                const uint32_t idx = l < kk ? l : l - kk;
                if (idx < (uint32_t)n_comp) meta[l][idx] = l < kk ? 1.0f : -1.0f;
            }
        };
        if (ncomp <= 2u * n_comp || !data) {
            simplex();
        } else if (ncomp > 2000) {
            for (uint32_t l = 0; l < ncomp; l++)
                for (int dd = 0; dd < n_comp; dd++)
                    meta[l][dd] = (float)(2.0 * u01(seed, 17, (uint64_t)l * n_comp + dd) - 1.0);
        } else {
            // centroids in feature space, dense exp(-d^2 / max d^2) affinity, spectral
            std::vector<Chunk> cch;
            std::vector<uint32_t> c0(ncomp + 1, 0);
            for (uint32_t c = 0; c < ncomp; c++) {
                c0[c] = (uint32_t)cch.size();
                for (uint32_t b = start[c]; b < start[c + 1]; b += kCenChunk)
                    cch.push_back({c, b, std::min(b + kCenChunk, start[c + 1])});
            }
            c0[ncomp] = (uint32_t)cch.size();
            DevBuf<Chunk> dch(cch.size());
            DevBuf<double> part(cch.size() * (size_t)d);
            MRS_HIP(hipMemcpyAsync(dch.p, cch.data(), sizeof(Chunk) * cch.size(), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_centroid_partials, dim3((unsigned)cch.size()), dim3(kWg), 0, s,
                               dch.p, dperm.p, data, d, part.p);
            MRS_CHECK_LAUNCH();
            std::vector<double> hp(cch.size() * (size_t)d);
            MRS_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
            MRS_HIP(hipStreamSynchronize(s));
            std::vector<double> cen((size_t)ncomp * d, 0.0);
            for (uint32_t c = 0; c < ncomp; c++) {
                for (uint32_t q = c0[c]; q < c0[c + 1]; q++)
                    for (int f = 0; f < d; f++) cen[(size_t)c * d + f] += hp[(size_t)q * d + f];
                for (int f = 0; f < d; f++) cen[(size_t)c * d + f] /= (double)size[c];
            }
            std::vector<double> W((size_t)ncomp * ncomp, 0.0);
            double maxd2 = 0;
            for (uint32_t a = 0; a < ncomp; a++)
                for (uint32_t b = a + 1; b < ncomp; b++) {
                    double d2 = 0;
                    for (int f = 0; f < d; f++) {
                        double df = cen[(size_t)a * d + f] - cen[(size_t)b * d + f];
                        d2 += df * df;
                    }
                    W[(size_t)a * ncomp + b] = W[(size_t)b * ncomp + a] = d2;
                    maxd2 = std::max(maxd2, d2);
                }
            const double scl = maxd2 > 1e-12 ? 1.0 / maxd2 : 1.0;
            for (uint32_t a = 0; a < ncomp; a++)
                for (uint32_t b = 0; b < ncomp; b++)
                    W[(size_t)a * ncomp + b] = a == b ? 0.0 : std::exp(-W[(size_t)a * ncomp + b] * scl);
            auto e = host_small_spectral(W, (int)ncomp, n_comp, seed);
            for (uint32_t l = 0; l < ncomp; l++) meta[l] = e[l];
        }
        float mx = 0.0f;
        for (auto& r : meta)
            for (float x : r) mx = std::max(mx, std::fabs(x));
        if (mx <= 1e-8f) {
            simplex();
            mx = 0.0f;
            for (auto& r : meta)
                for (float x : r) mx = std::max(mx, std::fabs(x));
        }
        if (mx > 1e-8f)
            for (auto& r : meta)
                for (float& x : r) x /= mx;
        // largest |raw| of each spectral component
        std::vector<float> ma(ncomp, 0.0f);
        if (nl) {
            DevBuf<float> amax(nch);
            hipLaunchKernelGGL(k_chunk_absmax, dim3(nch), dim3(kWg), 0, s, dchunks.p, raw.p,
                               n_comp, amax.p);
            MRS_CHECK_LAUNCH();
            std::vector<float> ha(nch);
            MRS_HIP(hipMemcpyAsync(ha.data(), amax.p, sizeof(float) * nch, hipMemcpyDeviceToHost, s));
            MRS_HIP(hipStreamSynchronize(s));
            for (uint32_t q = 0; q < nl; q++)
                for (uint32_t c2 = chunk0[q]; c2 < chunk0[q + 1]; c2++)
                    ma[lc_comp[q]] = std::max(ma[lc_comp[q]], ha[c2]);
        }
        // per component placement around its centroid (multi_component_init :439-484)
        std::vector<float> hmeta((size_t)ncomp * n_comp), dr(ncomp), expansion(ncomp, 1.0f);
        for (uint32_t c = 0; c < ncomp; c++) {
            float r = 0.0f;  // meta_data_range: half the largest centroid distance
            for (uint32_t o = 0; o < ncomp; o++) {
                float d2 = 0;
                for (int dd = 0; dd < n_comp; dd++) {
                    float df = meta[c][dd] - meta[o][dd];
                    d2 += df * df;
                }
                r = std::max(r, std::sqrt(d2));
            }
            dr[c] = r / 2.0f;
            if (spectral[c]) expansion[c] = ma[c] > 1e-8f ? dr[c] / ma[c] : 1.0f;
            for (int dd = 0; dd < n_comp; dd++) hmeta[(size_t)c * n_comp + dd] = meta[c][dd];
        }
        DevBuf<float> dmeta(hmeta.size()), ddr(ncomp), dexp(ncomp);
        DevBuf<uint8_t> dspec(ncomp);
        MRS_HIP(hipMemcpyAsync(dmeta.p, hmeta.data(), sizeof(float) * hmeta.size(), hipMemcpyHostToDevice, s));
        MRS_HIP(hipMemcpyAsync(ddr.p, dr.data(), sizeof(float) * ncomp, hipMemcpyHostToDevice, s));
        MRS_HIP(hipMemcpyAsync(dexp.p, expansion.data(), sizeof(float) * ncomp, hipMemcpyHostToDevice, s));
        MRS_HIP(hipMemcpyAsync(dspec.p, spectral.data(), ncomp, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_place, gn, dim3(kWg), 0, s, raw.p, dstart.p, ncomp, dperm.p, dmeta.p,
                           dexp.p, ddr.p, dspec.p, n, n_comp, seed, emb.p);
        MRS_CHECK_LAUNCH();
        // the host vectors above are read by the copies: finish before they go
        MRS_HIP(hipStreamSynchronize(s));
    }
    // centre, scale max |x| to the range (and, for a single component, add
    // the noise), written back in the caller's vertex order
    DevBuf<double> colp((size_t)kFinBlocks * n_comp);
    DevBuf<float> amaxp(kFinBlocks);
    hipLaunchKernelGGL(k_col_sums, dim3(kFinBlocks), dim3(kWg), 0, s, emb.p, n, n_comp, colp.p);
    hipLaunchKernelGGL(k_center_absmax, dim3(kFinBlocks), dim3(kWg), 0, s, emb.p, n, n_comp,
                       colp.p, amaxp.p);
    hipLaunchKernelGGL(k_finish_rows, gn, dim3(kWg), 0, s, emb.p, n, n_comp, amaxp.p, range,
                       noise, seed, dperm.p, out);
    MRS_CHECK_LAUNCH();
    MRS_HIP(hipStreamSynchronize(s));
}

}  // namespace mrs

extern "C" int mrs_spectral_layout(mrs_ctx* ctx, const mrs_graph* g, int32_t n_dim, uint64_t seed,
                                   float range, const float* data, int32_t d, float* out) {
    return mrs::guarded([&] {
        if (!ctx || !g || !out) throw mrs::Error(MRS_INVALID_ARGUMENT, "NULL argument");
        mrs::StreamScope scope(ctx->stream, ctx->pool);
        if (n_dim <= 0 || n_dim > 16) throw mrs::Error(MRS_INVALID_ARGUMENT, "n_dim must be 1..16");
        hipStream_t s = ctx->stream;
        mrs::DevView<float> xv(data, data ? (size_t)g->n * d : 0, s);
        mrs::DevView<float> ov(out, (size_t)g->n * n_dim, s, false);
        mrs::spectral_layout_dev(ctx, g, n_dim, seed, std::isnan(range) ? 10.0f : range, ov.p,
                                 xv.p, data ? d : 0);
        ov.download(s);
        MRS_HIP(hipStreamSynchronize(s));
    });
}

extern "C" int mrs_tridiag_eigen(const double* diag, const double* offdiag, int32_t m, double* evals,
                                 double* evecs) {
    return mrs::guarded([&] {
        if (m <= 0 || !diag || !evals || !evecs || (m > 1 && !offdiag))
            throw mrs::Error(MRS_INVALID_ARGUMENT, "bad tridiagonal problem");
        std::vector<double> a(diag, diag + m), b(offdiag ? offdiag : diag, offdiag ? offdiag + m - 1 : diag);
        std::vector<double> ev, vv;
        mrs::tridiag_eig(a, b, m, ev, vv);
        std::copy(ev.begin(), ev.end(), evals);
        std::copy(vv.begin(), vv.end(), evecs);
    });
}
