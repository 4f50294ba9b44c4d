// construct_umap_graph_gpu on the device (src/lib.rs:3376-3441):
//
//   smooth_knn_dist   (src/data/graph.rs:499-565)  one thread per row, the
//                     bandwidth binary search carried in FP64
//   knn_to_coo        (:585-662)                   fused into the same thread (f32)
//   symmetrise_graph  (:684-694 -> :387-468)       every kNN entry is emitted into
//                     its source row (out) and its target row (in); rows are
//                     sorted by (column, direction) with per-row bitonic sorts
//                     (registers / LDS / global for the rare long row) and
//                     merged by one wavefront per row into the fuzzy union
//   filter_weak_edges (:776-853)                   device max + threshold flags
//   UmapCsrGraph::from_graph (src/training/umap_optimiser_gpu.rs:132-209)
//                     edges are the kept upper-triangle entries in row-major
//                     order; a node's CSR slice lists its incident edges in edge
//                     order, which for a symmetric pattern is its sorted
//                     neighbour row; ranks come from three exclusive scans.
//
// The union keeps every (row, col) of out(row) U in(row) -- a pattern that is
// symmetric by construction -- so the mirror (col, row) of every entry comes
// from one stable sort of the columns, and no host adjacency list is ever built.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <vector>

#include "graph.hpp"
#include "umap_state.hpp"

namespace mrs {

namespace {

constexpr int kWg = 256;
constexpr uint64_t kInf64 = ~0ull;
constexpr int kMidCap = 2048;  // rows up to this length are sorted in LDS

// local_connectivity_distance, src/data/graph.rs:107-125 (f32)
__device__ inline float lc_distance(const float* d, int len, float lc) {
    if (lc <= 0.0f || len == 0) return 0.0f;
    float offset = lc - 1.0f;
    if (offset < 0.0f) offset = 0.0f;
    float whole = floorf(offset);
    float fraction = offset - whole;
    int idx = (int)whole;
    if (idx > len - 1) idx = len - 1;
    if (fraction > 0.0f && idx + 1 < len) return d[idx] * (1.0f - fraction) + d[idx + 1] * fraction;
    return d[idx];
}

// smooth_knn_dist (bandwidth search in FP64) + knn_to_coo weights (f32).
// kSmoothLanes lanes per row: lane l takes the row's neighbours l, l + 16, ...
// and a fixed xor tree joins their terms, so every lane of a row sees the
// same sum and takes the same branch of the search, and a wave's rows wait
// for the slowest of four rows rather than of sixty-four.
constexpr int kSmoothLanes = 16;
__global__ __launch_bounds__(kWg) void k_smooth_knn(const float* __restrict__ dist,
                                                    const int32_t* __restrict__ idx, uint32_t n,
                                                    int k, float lc, double tol, double target,
                                                    int n_iter, float* __restrict__ sigma,
                                                    float* __restrict__ rho,
                                                    float* __restrict__ w_out,
                                                    uint32_t* __restrict__ out_cnt,
                                                    uint32_t* __restrict__ in_cnt,
                                                    uint32_t* __restrict__ bad_idx) {
    const uint32_t i = (blockIdx.x * kWg + threadIdx.x) / kSmoothLanes;
    const int sub = threadIdx.x & (kSmoothLanes - 1);
    if (i >= n) return;  // whole row groups: the xor trees stay within a group
    const float* d = dist + (size_t)i * k;
    const float r = lc_distance(d, k, lc);
    const double sentinel = (double)FLT_MAX;
    double lo = 0.0, hi = sentinel, mid = 1.0;
    for (int it = 0; it < n_iter; it++) {
        // the reciprocal is taken once per pass and multiplied in, as the
        // reference does (graph.rs:538-540)
        const double inv_mid = 1.0 / mid;
        double s = 0.0;
        for (int j = sub; j < k; j += kSmoothLanes) {
            float a = d[j] - r;
            a = a > 0.0f ? a : 0.0f;
            s += exp(-(double)a * inv_mid);
        }
#pragma unroll
        for (int o = kSmoothLanes / 2; o >= 1; o >>= 1) s += __shfl_xor(s, o, kSmoothLanes);
        if (fabs(s - target) < tol) break;
        if (s > target) {
            hi = mid;
            mid = (lo + hi) / 2.0;
        } else {
            lo = mid;
            if (hi == sentinel) mid *= 2.0;
            else mid = (lo + hi) / 2.0;
        }
    }
    const float sg = (float)mid;
    if (sub == 0) {
        sigma[i] = sg;
        rho[i] = r;
    }
    const bool smooth = sg > 0.0f;
    const float inv_sigma = smooth ? 1.0f / sg : 0.0f;
    uint32_t cnt = 0;
    for (int j = sub; j < k; j += kSmoothLanes) {
        const int32_t c = idx[(size_t)i * k + j];
        float a = d[j] - r;
        a = a > 0.0f ? a : 0.0f;
        const float w = smooth ? expf(-a * inv_sigma) : (a > 0.0f ? 0.0f : 1.0f);
        w_out[(size_t)i * k + j] = w;
        if (c < 0 || (uint32_t)c >= n) {
            atomicAdd(bad_idx, 1u);  // the caller's precomputed kNN names no such point
        } else if ((uint32_t)c != i) {
            cnt++;
            atomicAdd(&in_cnt[c], 1u);
        }
    }
#pragma unroll
    for (int o = kSmoothLanes / 2; o >= 1; o >>= 1) cnt += __shfl_xor(cnt, o, kSmoothLanes);
    if (sub == 0) out_cnt[i] = cnt;
}

__global__ __launch_bounds__(kWg) void k_add_u32(const uint32_t* __restrict__ a,
                                                 const uint32_t* __restrict__ b,
                                                 uint32_t* __restrict__ out, uint32_t n) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i < n) out[i] = a[i] + b[i];
}

// Emit every kNN entry (i -> c, w) into row i (direction 0) and row c
// (direction 1).  Key = col << 33 | dir << 32 | bits(w): sorting the 64-bit
// key orders a row by (column, direction).
__global__ __launch_bounds__(kWg) void k_scatter_raw(const int32_t* __restrict__ idx,
                                                     const float* __restrict__ w, uint32_t n,
                                                     int k, const uint32_t* __restrict__ raw_off,
                                                     const uint32_t* __restrict__ out_cnt,
                                                     uint32_t* __restrict__ cursor,
                                                     uint64_t* __restrict__ raw) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i >= n) return;
    uint32_t pos = raw_off[i];
    for (int j = 0; j < k; j++) {
        const int32_t c = idx[(size_t)i * k + j];
        if (c < 0 || (uint32_t)c >= n || (uint32_t)c == i) continue;
        const uint32_t wb = __float_as_uint(w[(size_t)i * k + j]);
        raw[pos++] = ((uint64_t)(uint32_t)c << 33) | (uint64_t)wb;
        const uint32_t q = raw_off[c] + out_cnt[c] + atomicAdd(&cursor[c], 1u);
        raw[q] = ((uint64_t)i << 33) | (1ull << 32) | (uint64_t)wb;
    }
}

__device__ inline uint64_t shfl_xor_u64(uint64_t v, int mask) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor(lo, mask, 64);
    hi = __shfl_xor(hi, mask, 64);
    return ((uint64_t)hi << 32) | lo;
}

// One wave per row.  Rows of length <= 64 are sorted in registers (bitonic,
// "flip" form: every compare puts the minimum at the lower index, so the
// +inf padding never moves); longer rows are appended to the mid/long lists.
__global__ __launch_bounds__(kWg) void k_sort_rows_small(uint64_t* __restrict__ raw,
                                                         const uint32_t* __restrict__ raw_off,
                                                         uint32_t n, uint32_t* __restrict__ lists,
                                                         uint32_t* __restrict__ list_cnt) {
    const uint32_t row = blockIdx.x * (kWg / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const uint32_t base = raw_off[row];
    const uint32_t L = raw_off[row + 1] - base;
    if (L <= 1) return;
    if (L > 64) {
        if (lane == 0) {
            if (L <= (uint32_t)kMidCap) {
                uint32_t q = atomicAdd(&list_cnt[0], 1u);
                lists[q] = row;
            } else {
                uint32_t q = atomicAdd(&list_cnt[1], 1u);
                lists[n + q] = row;
            }
        }
        return;
    }
    uint64_t v = (uint32_t)lane < L ? raw[base + lane] : kInf64;
    for (int s = 2; s <= 64; s <<= 1) {
        uint64_t o = shfl_xor_u64(v, s - 1);
        bool lower = (lane & (s - 1)) < (s >> 1);
        v = lower ? (o < v ? o : v) : (o > v ? o : v);
        for (int j = s >> 2; j >= 1; j >>= 1) {
            o = shfl_xor_u64(v, j);
            lower = (lane & j) == 0;
            v = lower ? (o < v ? o : v) : (o > v ? o : v);
        }
    }
    if ((uint32_t)lane < L) raw[base + lane] = v;
}

// compare-exchange of a[i], a[p] (i < p): minimum to i; p >= L is +inf
__device__ inline void cmpx(uint64_t* a, uint32_t i, uint32_t p, uint32_t L) {
    if (p >= L) return;
    uint64_t x = a[i], y = a[p];
    if (y < x) {
        a[i] = y;
        a[p] = x;
    }
}

__device__ inline void bitonic_block(uint64_t* a, uint32_t L, uint32_t npad, int tid, int nthr) {
    for (uint32_t s = 2; s <= npad; s <<= 1) {
        for (uint32_t q = tid; q < npad / 2; q += nthr) {
            uint32_t b = q / (s >> 1), t = q % (s >> 1);
            cmpx(a, b * s + t, b * s + s - 1 - t, L);
        }
        __syncthreads();
        for (uint32_t j = s >> 2; j >= 1; j >>= 1) {
            for (uint32_t q = tid; q < npad / 2; q += nthr) {
                uint32_t i = (q / j) * 2 * j + (q % j);
                cmpx(a, i, i + j, L);
            }
            __syncthreads();
        }
    }
}

// 64 < L <= kMidCap: one single-wave block per row, sorted in LDS.
__global__ __launch_bounds__(64) void k_sort_rows_mid(uint64_t* __restrict__ raw,
                                                      const uint32_t* __restrict__ raw_off,
                                                      const uint32_t* __restrict__ rows) {
    __shared__ uint64_t sh[kMidCap];
    const uint32_t row = rows[blockIdx.x];
    const uint32_t base = raw_off[row];
    const uint32_t L = raw_off[row + 1] - base;
    uint32_t npad = 1;
    while (npad < L) npad <<= 1;
    for (uint32_t q = threadIdx.x; q < L; q += 64) sh[q] = raw[base + q];
    __syncthreads();
    bitonic_block(sh, L, npad, threadIdx.x, 64);
    for (uint32_t q = threadIdx.x; q < L; q += 64) raw[base + q] = sh[q];
}

// L > kMidCap (rare hubs): one 256-thread block per row, sorted in global memory.
__global__ __launch_bounds__(kWg) void k_sort_rows_long(uint64_t* __restrict__ raw,
                                                        const uint32_t* __restrict__ raw_off,
                                                        const uint32_t* __restrict__ rows) {
    const uint32_t row = rows[blockIdx.x];
    const uint32_t base = raw_off[row];
    const uint32_t L = raw_off[row + 1] - base;
    uint32_t npad = 1;
    while (npad < L) npad <<= 1;
    bitonic_block(raw + base, L, npad, threadIdx.x, kWg);
}

// merge_row / union_len (src/data/graph.rs:278-363) over a sorted raw row,
// one wavefront per row.  An entry heads its column's group when the entry
// before it has another column; the heads are counted with ballots, and in the
// write pass each head folds its group (the last entry of each direction
// wins, as in a serial walk of the sorted row) and writes (row, col,
// combine(w_ij, w_ji)) at the row's offset plus the heads before it.
// Pass 0 counts the distinct columns of each row.
template <bool WRITE>
__global__ __launch_bounds__(kWg) void k_union_rows(
    const uint64_t* __restrict__ raw, const uint32_t* __restrict__ raw_off, uint32_t n,
    uint32_t* __restrict__ ucnt, const uint32_t* __restrict__ u_off, uint32_t* __restrict__ u_row,
    uint32_t* __restrict__ u_col, float* __restrict__ u_w, float mix, float directed) {
    const uint32_t row = blockIdx.x * (kWg / 64) + (threadIdx.x >> 6);  // uniform per wave
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t beg = raw_off[row], end = raw_off[row + 1];
    const uint32_t o = WRITE ? u_off[row] : 0u;
    uint32_t cnt = 0;
    for (uint32_t c0 = beg; c0 < end; c0 += 64) {
        const uint32_t p = c0 + (uint32_t)lane;
        bool head = false;
        uint32_t col = 0;
        if (p < end) {
            col = (uint32_t)(raw[p] >> 33);
            head = p == beg || (uint32_t)(raw[p - 1] >> 33) != col;
        }
        const uint64_t heads = __ballot(head);
        if (WRITE && head) {
            float wij = 0.0f, wji = 0.0f;
            for (uint32_t q = p; q < end; q++) {
                const uint64_t e = raw[q];
                if ((uint32_t)(e >> 33) != col) break;
                const float w = __uint_as_float((uint32_t)e);
                if ((e >> 32) & 1ull) wji = w;
                else wij = w;
            }
            // f32, unfused: union = w_ij + w_ji - w_ij*w_ji; mix*union + (1-mix)*w_ij
            const float sum = __fadd_rn(wij, wji);
            const float uni = __fsub_rn(sum, __fmul_rn(wij, wji));
            const float wsym = __fadd_rn(__fmul_rn(mix, uni), __fmul_rn(directed, wij));
            const uint32_t at = o + cnt + (uint32_t)__popcll(heads & below);
            u_row[at] = row;
            u_col[at] = col;
            u_w[at] = wsym;
        }
        cnt += (uint32_t)__popcll(heads);
    }
    if (!WRITE && lane == 0) ucnt[row] = cnt;
}

__global__ __launch_bounds__(kWg) void k_max_pos(const float* __restrict__ w, uint64_t m,
                                                 uint32_t* __restrict__ max_bits) {
    uint32_t best = 0;
    for (uint64_t p = (uint64_t)blockIdx.x * kWg + threadIdx.x; p < m;
         p += (uint64_t)gridDim.x * kWg) {
        float x = w[p];
        if (x > 0.0f) {
            uint32_t b = __float_as_uint(x);
            best = b > best ? b : best;
        }
    }
    atomicMax(max_bits, best);
}

// filter_weak_edges flags and the flag vectors for the scans.  The mirror of
// entry p = (r, c) is the position of (c, r).  The union is grouped by row with
// columns ascending and its pattern is symmetric, so the entries taken in
// (column, row) order are the same (row, column) list: a stable sort of the
// columns (values = entry index) puts p at position mirror[p], and the sorted
// columns equal u_row element by element.  That equality is checked here
// (*bad is raised if the pattern is not symmetric).
__global__ __launch_bounds__(kWg) void k_flags(
    const uint32_t* __restrict__ u_row, const uint32_t* __restrict__ u_col,
    const float* __restrict__ u_w, const uint32_t* __restrict__ sorted_col, uint64_t m,
    const uint32_t* __restrict__ max_bits, float n_epochs_f, uint32_t* __restrict__ keep,
    uint32_t* __restrict__ upper, uint32_t* __restrict__ bad) {
    uint64_t p = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (p >= m) return;
    const float thr = __uint_as_float(*max_bits) / n_epochs_f;
    const float w = u_w[p];
    const uint32_t r = u_row[p], c = u_col[p];
    const uint32_t kp = (w > 0.0f && w >= thr) ? 1u : 0u;
    keep[p] = kp;
    upper[p] = (kp && c > r) ? 1u : 0u;
    if (sorted_col[p] != r) *bad = 1u;
}

__global__ __launch_bounds__(kWg) void k_present(const uint32_t* __restrict__ u_row,
                                                 const uint32_t* __restrict__ u_col,
                                                 const uint32_t* __restrict__ keep,
                                                 const uint32_t* __restrict__ mirror, uint64_t m,
                                                 uint32_t* __restrict__ present) {
    uint64_t p = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (p >= m) return;
    const uint32_t r = u_row[p], c = u_col[p];
    present[p] = c > r ? keep[p] : (c < r ? keep[mirror[p]] : 0u);
}

__global__ __launch_bounds__(kWg) void k_offsets(const uint32_t* __restrict__ u_off,
                                                 const uint32_t* __restrict__ present_scan,
                                                 uint32_t total, uint64_t m, uint32_t n,
                                                 uint32_t* __restrict__ off) {
    uint32_t v = blockIdx.x * kWg + threadIdx.x;
    if (v > n) return;
    uint32_t q = u_off[v];
    off[v] = q < m ? present_scan[q] : total;
}

__global__ __launch_bounds__(kWg) void k_emit(
    const uint32_t* __restrict__ u_row, const uint32_t* __restrict__ u_col,
    const float* __restrict__ u_w, const uint32_t* __restrict__ keep,
    const uint32_t* __restrict__ mirror, const uint32_t* __restrict__ upper,
    const uint32_t* __restrict__ keep_scan, const uint32_t* __restrict__ up_scan,
    const uint32_t* __restrict__ present_scan, const uint32_t* __restrict__ off, uint64_t m,
    uint32_t* __restrict__ coo_row, uint32_t* __restrict__ coo_col, float* __restrict__ coo_val,
    uint32_t* __restrict__ csr_other, uint32_t* __restrict__ csr_edge_idx,
    uint4* __restrict__ edges, float* __restrict__ edge_w) {
    uint64_t p = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (p >= m) return;
    const uint32_t r = u_row[p], c = u_col[p];
    if (keep[p]) {
        const uint32_t q = keep_scan[p];
        coo_row[q] = r;
        coo_col[q] = c;
        coo_val[q] = u_w[p];
    }
    const uint32_t mp = mirror[p];
    const bool pres = c > r ? keep[p] != 0 : (c < r ? keep[mp] != 0 : false);
    if (pres) {
        const uint32_t q = present_scan[p];
        csr_other[q] = c;
        csr_edge_idx[q] = c > r ? up_scan[p] : up_scan[mp];
    }
    if (upper[p]) {
        const uint32_t e = up_scan[p];
        edge_w[e] = u_w[p];
        // the edge's CSR slots under r and c: the edge_local keys of its negatives
        edges[e] = make_uint4(r, c, present_scan[p] - off[r], present_scan[mp] - off[c]);
    }
}

// epochs_per_sample = max_w / w (1e8 when the normalised weight is 0),
// UmapCsrGraph::from_graph, umap_optimiser_gpu.rs:181-198
__global__ __launch_bounds__(kWg) void k_periods(const float* __restrict__ w, uint64_t ne,
                                                 const uint32_t* __restrict__ max_bits,
                                                 float* __restrict__ per) {
    uint64_t e = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (e >= ne) return;
    const float mx = __uint_as_float(*max_bits);
    const float norm = w[e] / mx;
    per[e] = norm > 0.0f ? 1.0f / norm : 1e8f;
}

// ------------------------------------------------------------ DensState

// edge_distances_from_knn + the row-wise f64 accumulation of original_log_radii
// (src/utils/density.rs:309-398).  kDensLanes lanes per row: lane l takes the
// row's entries l, l + kDensLanes, ... and the lanes' f64 sums meet in a fixed
// xor tree, so the result does not depend on scheduling.  An entry's distance
// is looked up in row i's kNN list, else in row j's.
constexpr int kDensLanes = 16;
constexpr int kDensStage = 128;
__global__ __launch_bounds__(kWg) void k_dens_rows(
    const uint32_t* __restrict__ u_off, const uint32_t* __restrict__ u_col,
    const float* __restrict__ u_w, const uint32_t* __restrict__ keep,
    const int32_t* __restrict__ knn_idx, const float* __restrict__ knn_dist, int k, uint32_t n,
    int squared, double* __restrict__ ro, float* __restrict__ mu_sum) {
    // row i's kNN indices staged in LDS (k <= kDensStage), where the row
    // group's lanes scan them for every entry (a broadcast read each)
    __shared__ int32_t own[kWg / kDensLanes][kDensStage];
    const uint32_t i = (blockIdx.x * kWg + threadIdx.x) / kDensLanes;
    const uint32_t sub = threadIdx.x & (kDensLanes - 1);
    const uint32_t grp = threadIdx.x / kDensLanes;
    const bool staged = k <= kDensStage;
    if (staged && i < n)
        for (int t = (int)sub; t < k; t += kDensLanes) own[grp][t] = knn_idx[(size_t)i * k + t];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();  // a row group lies within one wave
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double acc = 0.0, ms = 0.0;
    if (i < n) {
        for (uint32_t p = u_off[i] + sub; p < u_off[i + 1]; p += kDensLanes) {
            if (!keep[p]) continue;
            const uint32_t j = u_col[p];
            float raw = 0.0f;
            bool found = false;
            for (int t = 0; t < k; t++)
                if ((uint32_t)(staged ? own[grp][t] : knn_idx[(size_t)i * k + t]) == j) {
                    raw = knn_dist[(size_t)i * k + t];
                    found = true;
                    break;
                }
            if (!found)
                for (int t = 0; t < k; t++)
                    if ((uint32_t)knn_idx[(size_t)j * k + t] == i) {
                        raw = knn_dist[(size_t)j * k + t];
                        break;
                    }
            const float dsq = squared ? raw : raw * raw;
            const double mu = (double)u_w[p];
            acc += mu * (double)dsq;
            ms += mu;
        }
    }
#pragma unroll
    for (int o = kDensLanes / 2; o >= 1; o >>= 1) {
        acc += __shfl_xor(acc, o, kDensLanes);
        ms += __shfl_xor(ms, o, kDensLanes);
    }
    if (i < n && sub == 0) {
        const double den = ms > DBL_EPSILON ? ms : DBL_EPSILON;
        ro[i] = log(1e-8 + acc / den);
        mu_sum[i] = (float)ms;
    }
}

constexpr int kRedBlocks = 256;

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

// fixed-grid f64 partial sums: mode 0 sum(ro), 1 sum((ro-mean)^2), 2 sum((double)mu_sum)
__global__ __launch_bounds__(kWg) void k_dens_reduce(const double* __restrict__ ro,
                                                     const float* __restrict__ mu_sum, uint32_t n,
                                                     int mode, const double* __restrict__ mean_p,
                                                     double* __restrict__ partials) {
    __shared__ double sh[kWg];
    const double mean = mode == 1 ? *mean_p : 0.0;
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * kWg + threadIdx.x; i < n; i += kRedBlocks * kWg) {
        if (mode == 0) acc += ro[i];
        else if (mode == 1) {
            double d = ro[i] - mean;
            acc += d * d;
        } else acc += (double)mu_sum[i];
    }
    double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void k_finish_sum(const double* __restrict__ partials, double scale,
                             double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double t = 0.0;
    for (int q = 0; q < kRedBlocks; q++) t += partials[q];
    *out = t * scale;
}

__global__ __launch_bounds__(kWg) void k_standardise(const double* __restrict__ ro, uint32_t n,
                                                     double mean, double sd,
                                                     float* __restrict__ r) {
    uint32_t i = blockIdx.x * kWg + threadIdx.x;
    if (i < n) r[i] = (float)((ro[i] - mean) / sd);
}

uint32_t read_u32(const uint32_t* p, hipStream_t s) {
    uint32_t v = 0;
    MRS_HIP(hipMemcpyAsync(&v, p, sizeof(v), hipMemcpyDeviceToHost, s));
    MRS_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace

// filter_weak_edges on the fuzzy union in g->u_* (m entries, rows grouped,
// cols ascending, symmetric pattern), then the filtered COO, the CSR under both
// endpoints and the per-edge arrays of the epoch kernels.
void finish_graph(mrs_graph* g, uint32_t m) {
    const uint32_t n = (uint32_t)g->n;
    const int32_t n_epochs = g->n_epochs;
    hipStream_t s = g->ctx->stream;
    const size_t ma = m ? m : 1;
    // filter flags
    DevBuf<uint32_t> max_bits(2);
    MRS_HIP(hipMemsetAsync(max_bits.p, 0, max_bits.bytes(), s));
    if (m) {
        hipLaunchKernelGGL(k_max_pos, dim3(std::min<unsigned>(div_up(m, kWg), 1024)),
                           dim3(kWg), 0, s, g->u_w.p, (uint64_t)m, max_bits.p);
        MRS_CHECK_LAUNCH();
    }
    g->u_keep.alloc(ma);
    DevBuf<uint32_t> mirror(ma), upper(ma), present(ma), keep_scan(ma), up_scan(ma),
        present_scan(ma);
    if (m) {
        // mirrors: a stable sort of the columns carrying the entry index
        DevBuf<uint32_t> sorted_col(m), asym(1);
        radix_sort_pairs_u32(g->u_col.p, nullptr, m, n > 1 ? 32 - __builtin_clz(n - 1) : 1,
                             sorted_col.p, mirror.p, s);
        MRS_HIP(hipMemsetAsync(asym.p, 0, sizeof(uint32_t), s));
        hipLaunchKernelGGL(k_flags, dim3(div_up(m, kWg)), dim3(kWg), 0, s, g->u_row.p,
                           g->u_col.p, g->u_w.p, sorted_col.p, (uint64_t)m, max_bits.p,
                           (float)n_epochs, g->u_keep.p, upper.p, asym.p);
        MRS_CHECK_LAUNCH();
        if (read_u32(asym.p, s))
            throw Error(MRS_DEVICE_ERROR, "fuzzy union pattern is not symmetric");
        hipLaunchKernelGGL(k_present, dim3(div_up(m, kWg)), dim3(kWg), 0, s, g->u_row.p,
                           g->u_col.p, g->u_keep.p, mirror.p, (uint64_t)m, present.p);
        MRS_CHECK_LAUNCH();
    }
    DevBuf<uint32_t> totals(3);
    exclusive_scan_u32(g->u_keep.p, keep_scan.p, m, totals.p + 0, s);
    exclusive_scan_u32(upper.p, up_scan.p, m, totals.p + 1, s);
    exclusive_scan_u32(present.p, present_scan.p, m, totals.p + 2, s);
    uint32_t tot[3];
    MRS_HIP(hipMemcpyAsync(tot, totals.p, sizeof(tot), hipMemcpyDeviceToHost, s));
    MRS_HIP(hipStreamSynchronize(s));
    // the totals size every output array below: check them against the union
    if (tot[0] > m || tot[1] > m || tot[2] > m)
        throw Error(MRS_DEVICE_ERROR, "graph scan totals exceed the union size");
    g->nnz = tot[0];
    g->n_edges = tot[1];
    if ((uint64_t)tot[2] != 2ull * tot[1])
        throw Error(MRS_DEVICE_ERROR, "graph CSR size mismatch (2E != present entries)");
    const size_t nn = g->nnz ? g->nnz : 1, ne = g->n_edges ? g->n_edges : 1;
    g->coo_row.alloc(nn);
    g->coo_col.alloc(nn);
    g->coo_val.alloc(nn);
    g->off.alloc((size_t)n + 1);
    g->csr_other.alloc(2 * ne);
    g->csr_edge_idx.alloc(2 * ne);
    g->edges.alloc(ne);
    g->edge_w.alloc(ne);
    g->per.alloc(ne);
    hipLaunchKernelGGL(k_offsets, dim3(div_up((uint64_t)n + 1, kWg)), dim3(kWg), 0, s,
                       g->u_off.p, present_scan.p, tot[2], (uint64_t)m, n, g->off.p);
    MRS_CHECK_LAUNCH();
    if (m) {
        hipLaunchKernelGGL(k_emit, dim3(div_up(m, kWg)), dim3(kWg), 0, s, g->u_row.p,
                           g->u_col.p, g->u_w.p, g->u_keep.p, mirror.p, upper.p, keep_scan.p,
                           up_scan.p, present_scan.p, g->off.p, (uint64_t)m, g->coo_row.p,
                           g->coo_col.p, g->coo_val.p, g->csr_other.p, g->csr_edge_idx.p,
                           g->edges.p, g->edge_w.p);
        MRS_CHECK_LAUNCH();
    }
    if (g->n_edges) {
        hipLaunchKernelGGL(k_max_pos, dim3(std::min<unsigned>(div_up(g->n_edges, kWg), 1024)),
                           dim3(kWg), 0, s, g->edge_w.p, (uint64_t)g->n_edges,
                           max_bits.p + 1);
        MRS_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_periods, dim3(div_up(g->n_edges, kWg)), dim3(kWg), 0, s,
                           g->edge_w.p, (uint64_t)g->n_edges, max_bits.p + 1, g->per.p);
        MRS_CHECK_LAUNCH();
    }
}

// ---- import of a CoordinateList (src/data/structures.rs:22-31)

// keys of the caller's entries, with the checks: indices in [0, n), no
// diagonal (bad = 1)
__global__ __launch_bounds__(kWg) void k_coo_keys(const int32_t* __restrict__ rows,
                                                  const int32_t* __restrict__ cols, uint64_t m,
                                                  uint32_t n, uint32_t* __restrict__ krow,
                                                  uint32_t* __restrict__ kcol,
                                                  uint32_t* __restrict__ bad) {
    const uint64_t p = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (p >= m) return;
    const int32_t r = rows[p], c = cols[p];
    if (r < 0 || c < 0 || (uint32_t)r >= n || (uint32_t)c >= n || r == c) {
        *bad = 1u;
        krow[p] = 0u;
        kcol[p] = 0u;
        return;
    }
    krow[p] = (uint32_t)r;
    kcol[p] = (uint32_t)c;
}

// out[q] = in[idx[q]]
__global__ __launch_bounds__(kWg) void k_gather_u32(const uint32_t* __restrict__ in,
                                                    const uint32_t* __restrict__ idx, uint64_t m,
                                                    uint32_t* __restrict__ out) {
    const uint64_t q = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (q < m) out[q] = in[idx[q]];
}

// The union arrays in (row, col) order; a repeated (row, col) sets bad = 2.
__global__ __launch_bounds__(kWg) void k_coo_union(const uint32_t* __restrict__ srow,
                                                   const uint32_t* __restrict__ idx,
                                                   const uint32_t* __restrict__ kcol,
                                                   const float* __restrict__ vals, uint64_t m,
                                                   uint32_t* __restrict__ u_row,
                                                   uint32_t* __restrict__ u_col,
                                                   float* __restrict__ u_w,
                                                   uint32_t* __restrict__ bad) {
    const uint64_t q = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (q >= m) return;
    const uint32_t e = idx[q];
    u_row[q] = srow[q];
    u_col[q] = kcol[e];
    u_w[q] = vals[e];
    if (q > 0) {
        const uint32_t e0 = idx[q - 1];
        if (srow[q - 1] == srow[q] && kcol[e0] == kcol[e]) *bad = 2u;
    }
}

// off[v] = first entry of row v in the row-sorted list (off[n] = m)
__global__ __launch_bounds__(kWg) void k_sorted_row_offsets(const uint32_t* __restrict__ row,
                                                            uint64_t m, uint32_t n,
                                                            uint32_t* __restrict__ off) {
    const uint32_t v = blockIdx.x * kWg + threadIdx.x;
    if (v > n) return;
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (row[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    off[v] = (uint32_t)lo;
}

__device__ inline uint32_t find_in_row(const uint32_t* __restrict__ u_col, uint32_t lo,
                                       uint32_t hi, uint32_t key) {
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (u_col[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// every (r, c) needs its mirror (c, r): bad = 3 otherwise
__global__ __launch_bounds__(kWg) void k_coo_symmetric(const uint32_t* __restrict__ u_row,
                                                       const uint32_t* __restrict__ u_col,
                                                       const uint32_t* __restrict__ u_off,
                                                       uint64_t m, uint32_t* __restrict__ bad) {
    const uint64_t p = (uint64_t)blockIdx.x * kWg + threadIdx.x;
    if (p >= m) return;
    const uint32_t r = u_row[p], c = u_col[p];
    const uint32_t q = find_in_row(u_col, u_off[c], u_off[c + 1], r);
    if (!(q < u_off[c + 1] && u_col[q] == r)) *bad = 3u;
}

// A graph from a caller's CoordinateList: the symmetric fuzzy graph that
// symmetrise_graph (+ filter_weak_edges) produces, in any entry order.  The
// entries are sorted by (row, col) and then go through the same filter and
// layout as a graph built from kNN (re-filtering an already filtered graph
// with its own n_epochs keeps it as it is).  The graph has no kNN behind it:
// no sigma/rho and no densMAP state.
mrs_graph* graph_from_coo(mrs_ctx* ctx, int64_t n64, const int32_t* rows, const int32_t* cols,
                          const float* vals, int64_t m64, int32_t n_epochs) {
    if (n64 <= 0) throw Error(MRS_NO_DATA, "UMAP: empty graph (NoData)");
    if (n64 > (int64_t)UINT32_MAX - 1 || m64 < 0 || m64 >= (int64_t)UINT32_MAX)
        throw Error(MRS_INVALID_ARGUMENT, "graph too large for 32-bit indices");
    if (n_epochs <= 0) throw Error(MRS_INVALID_ARGUMENT, "n_epochs must be positive");
    const uint32_t n = (uint32_t)n64, m = (uint32_t)m64;
    hipStream_t s = ctx->stream;
    auto* g = new mrs_graph();
    try {
        g->ctx = ctx;
        g->n = n64;
        g->k = 0;
        g->n_epochs = n_epochs;
        g->n_union = m;
        const size_t ma = m ? m : 1;
        g->u_off.alloc((size_t)n + 1);
        g->u_row.alloc(ma);
        g->u_col.alloc(ma);
        g->u_w.alloc(ma);
        DevBuf<uint32_t> bad(1);
        MRS_HIP(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), s));
        if (m) {
            DevView<int32_t> rv(rows, m, s), cv(cols, m, s);
            DevView<float> vv(vals, m, s);
            DevBuf<uint32_t> krow(m), kcol(m), t_key(m), idx1(m), row1(m), srow(m), idx2(m);
            const dim3 gm(div_up(m, kWg));
            hipLaunchKernelGGL(k_coo_keys, gm, dim3(kWg), 0, s, rv.p, cv.p, (uint64_t)m, n,
                               krow.p, kcol.p, bad.p);
            MRS_CHECK_LAUNCH();
            const int bits = n > 1 ? 32 - __builtin_clz(n - 1) : 1;
            // stable by column, then stable by row: (row, col) order
            radix_sort_pairs_u32(kcol.p, nullptr, m, bits, t_key.p, idx1.p, s);
            hipLaunchKernelGGL(k_gather_u32, gm, dim3(kWg), 0, s, krow.p, idx1.p, (uint64_t)m,
                               row1.p);
            MRS_CHECK_LAUNCH();
            radix_sort_pairs_u32(row1.p, idx1.p, m, bits, srow.p, idx2.p, s);
            hipLaunchKernelGGL(k_coo_union, gm, dim3(kWg), 0, s, srow.p, idx2.p, kcol.p, vv.p,
                               (uint64_t)m, g->u_row.p, g->u_col.p, g->u_w.p, bad.p);
            MRS_CHECK_LAUNCH();
        }
        hipLaunchKernelGGL(k_sorted_row_offsets, dim3(div_up((uint64_t)n + 1, kWg)), dim3(kWg), 0,
                           s, g->u_row.p, (uint64_t)m, n, g->u_off.p);
        MRS_CHECK_LAUNCH();
        uint32_t b = read_u32(bad.p, s);
        if (b == 0 && m) {
            hipLaunchKernelGGL(k_coo_symmetric, dim3(div_up(m, kWg)), dim3(kWg), 0, s, g->u_row.p,
                               g->u_col.p, g->u_off.p, (uint64_t)m, bad.p);
            MRS_CHECK_LAUNCH();
            b = read_u32(bad.p, s);
        }
        if (b == 1)
            throw Error(MRS_INVALID_ARGUMENT, "graph entry out of range or on the diagonal");
        if (b == 2) throw Error(MRS_INVALID_ARGUMENT, "graph entry (row, col) repeated");
        if (b == 3)
            throw Error(MRS_INVALID_ARGUMENT,
                        "graph is not symmetric: an entry (r, c) has no (c, r)");
        finish_graph(g, m);
    } catch (...) {
        delete g;
        throw;
    }
    return g;
}

mrs_graph* build_graph(mrs_ctx* ctx, const int32_t* knn_idx, const float* knn_dist, int64_t n64,
                       int32_t k, float bandwidth, float local_connectivity, float mix_weight,
                       int32_t n_epochs) {
    if (n64 <= 0) throw Error(MRS_NO_DATA, "UMAP: empty kNN input (NoData)");
    if (k <= 0) throw Error(MRS_INVALID_ARGUMENT, "k must be positive");
    if ((uint64_t)n64 * (uint64_t)k * 2ull >= (1ull << 32))
        throw Error(MRS_INVALID_ARGUMENT, "n*k too large for 32-bit graph indices");
    if (n_epochs <= 0) throw Error(MRS_INVALID_ARGUMENT, "n_epochs must be positive");
    const uint32_t n = (uint32_t)n64;
    hipStream_t s = ctx->stream;
    auto* g = new mrs_graph();
    try {
        g->ctx = ctx;
        g->n = n64;
        g->k = k;
        g->n_epochs = n_epochs;
        g->sigma.alloc(n);
        g->rho.alloc(n);
        DevBuf<float> w((size_t)n * k);
        DevBuf<uint32_t> out_cnt(n), in_cnt(n), raw_cnt(n + 1), raw_off(n + 1), cursor(n),
            bad_idx(1);
        MRS_HIP(hipMemsetAsync(bad_idx.p, 0, bad_idx.bytes(), s));
        MRS_HIP(hipMemsetAsync(in_cnt.p, 0, in_cnt.bytes(), s));
        MRS_HIP(hipMemsetAsync(cursor.p, 0, cursor.bytes(), s));
        MRS_HIP(hipMemsetAsync(raw_cnt.p, 0, raw_cnt.bytes(), s));
        hipLaunchKernelGGL(k_smooth_knn, dim3(div_up((uint64_t)n * kSmoothLanes, kWg)), dim3(kWg), 0,
                           s, knn_dist, knn_idx,
                           n, k, local_connectivity, (double)bandwidth, std::log((double)k), 64,
                           g->sigma.p, g->rho.p, w.p, out_cnt.p, in_cnt.p, bad_idx.p);
        MRS_CHECK_LAUNCH();
        // the reference indexes its point arrays with every neighbour id and
        // stops on one out of range; so does this build, before any use of it
        if (const uint32_t nbad = read_u32(bad_idx.p, s))
            throw Error(MRS_INVALID_ARGUMENT, std::to_string(nbad) +
                                                  " kNN indices outside [0, n)");
        hipLaunchKernelGGL(k_add_u32, dim3(div_up(n, kWg)), dim3(kWg), 0, s, out_cnt.p, in_cnt.p,
                           raw_cnt.p, n);
        MRS_CHECK_LAUNCH();
        exclusive_scan_u32(raw_cnt.p, raw_off.p, (uint64_t)n + 1, nullptr, s);
        const uint32_t n_raw = read_u32(raw_off.p + n, s);
        DevBuf<uint64_t> raw(n_raw ? n_raw : 1);
        hipLaunchKernelGGL(k_scatter_raw, dim3(div_up(n, kWg)), dim3(kWg), 0, s, knn_idx, w.p, n,
                           k, raw_off.p, out_cnt.p, cursor.p, raw.p);
        MRS_CHECK_LAUNCH();
        // per-row sorts
        DevBuf<uint32_t> lists(2 * (size_t)n), list_cnt(2);
        MRS_HIP(hipMemsetAsync(list_cnt.p, 0, list_cnt.bytes(), s));
        hipLaunchKernelGGL(k_sort_rows_small, dim3(div_up(n, kWg / 64)), dim3(kWg), 0, s, raw.p,
                           raw_off.p, n, lists.p, list_cnt.p);
        MRS_CHECK_LAUNCH();
        uint32_t cnts[2];
        MRS_HIP(hipMemcpyAsync(cnts, list_cnt.p, sizeof(cnts), hipMemcpyDeviceToHost, s));
        MRS_HIP(hipStreamSynchronize(s));
        if (cnts[0]) {
            hipLaunchKernelGGL(k_sort_rows_mid, dim3(cnts[0]), dim3(64), 0, s, raw.p, raw_off.p,
                               lists.p);
            MRS_CHECK_LAUNCH();
        }
        if (cnts[1]) {
            hipLaunchKernelGGL(k_sort_rows_long, dim3(cnts[1]), dim3(kWg), 0, s, raw.p, raw_off.p,
                               lists.p + n);
            MRS_CHECK_LAUNCH();
        }
        // fuzzy union
        DevBuf<uint32_t> ucnt(n + 1);
        MRS_HIP(hipMemsetAsync(ucnt.p, 0, ucnt.bytes(), s));
        hipLaunchKernelGGL(k_union_rows<false>, dim3(div_up(n, kWg / 64)), dim3(kWg), 0, s, raw.p,
                           raw_off.p, n, ucnt.p, nullptr, nullptr, nullptr, nullptr, mix_weight,
                           1.0f - mix_weight);
        MRS_CHECK_LAUNCH();
        g->u_off.alloc((size_t)n + 1);
        exclusive_scan_u32(ucnt.p, g->u_off.p, (uint64_t)n + 1, nullptr, s);
        const uint32_t m = read_u32(g->u_off.p + n, s);
        g->n_union = m;
        const size_t ma = m ? m : 1;
        g->u_row.alloc(ma);
        g->u_col.alloc(ma);
        g->u_w.alloc(ma);
        hipLaunchKernelGGL(k_union_rows<true>, dim3(div_up(n, kWg / 64)), dim3(kWg), 0, s, raw.p,
                           raw_off.p, n, nullptr, g->u_off.p, g->u_row.p, g->u_col.p, g->u_w.p,
                           mix_weight, 1.0f - mix_weight);
        MRS_CHECK_LAUNCH();
        raw.release();  // the union kernels that read it are queued before any reuse
        finish_graph(g, m);
    } catch (...) {
        delete g;
        throw;
    }
    return g;
}

float dens_state_on_device(mrs_graph* g, const int32_t* knn_idx, const float* knn_dist,
                           bool distances_squared, float* r_dev, float* mu_sum_dev) {
    const uint32_t n = (uint32_t)g->n;
    if (n == 0) throw Error(MRS_NO_DATA, "densMAP: empty graph (NoData)");
    hipStream_t s = g->ctx->stream;
    DevBuf<double> ro(n), partials(kRedBlocks), scal(3);
    hipLaunchKernelGGL(k_dens_rows, dim3(div_up((uint64_t)n * kDensLanes, kWg)), dim3(kWg), 0, s, g->u_off.p,
                       g->u_col.p, g->u_w.p, g->u_keep.p, knn_idx, knn_dist, g->k, n,
                       distances_squared ? 1 : 0, ro.p, mu_sum_dev);
    MRS_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_dens_reduce, dim3(kRedBlocks), dim3(kWg), 0, s, ro.p, mu_sum_dev, n, 0,
                       nullptr, partials.p);
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1), 0, s, partials.p, 1.0 / (double)n,
                       scal.p + 0);
    hipLaunchKernelGGL(k_dens_reduce, dim3(kRedBlocks), dim3(kWg), 0, s, ro.p, mu_sum_dev, n, 1,
                       scal.p + 0, partials.p);
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1), 0, s, partials.p,
                       n > 1 ? 1.0 / (double)(n - 1) : 0.0, scal.p + 1);
    hipLaunchKernelGGL(k_dens_reduce, dim3(kRedBlocks), dim3(kWg), 0, s, ro.p, mu_sum_dev, n, 2,
                       nullptr, partials.p);
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1), 0, s, partials.p, 1.0, scal.p + 2);
    MRS_CHECK_LAUNCH();
    double h[3];
    MRS_HIP(hipMemcpyAsync(h, scal.p, sizeof(h), hipMemcpyDeviceToHost, s));
    MRS_HIP(hipStreamSynchronize(s));
    // standardise, src/utils/density.rs:410-438
    const double sd = std::sqrt(h[1]);
    if (n < 2 || std::isnan(sd) || sd <= DBL_EPSILON)
        throw Error(MRS_DEGENERATE_LOCAL_RADII,
                    "densMAP: the local radii have no spread (DegenerateLocalRadii)");
    hipLaunchKernelGGL(k_standardise, dim3(div_up(n, kWg)), dim3(kWg), 0, s, ro.p, n, h[0], sd,
                       r_dev);
    MRS_CHECK_LAUNCH();
    MRS_HIP(hipStreamSynchronize(s));
    return (float)h[2];
}

}  // namespace mrs

// ------------------------------------------------------------------ C ABI

using mrs::Error;

extern "C" {

int mrs_fuzzy_graph(mrs_ctx* ctx, const int32_t* knn_idx, const float* knn_dist, int64_t n,
                    int32_t k, const mrs_umap_graph_params* gp, int32_t n_epochs,
                    mrs_graph** out) {
    return mrs::guarded([&] {
        if (!ctx || !out || !gp) throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        mrs::StreamScope scope(ctx->stream, ctx->pool);
        if (n <= 0 || !knn_idx || !knn_dist) throw Error(MRS_NO_DATA, "UMAP: empty kNN (NoData)");
        MRS_HIP(hipSetDevice(ctx->device));
        const size_t nk = (size_t)n * (size_t)k;
        mrs::DevView<int32_t> idx(knn_idx, nk, ctx->stream);
        mrs::DevView<float> dist(knn_dist, nk, ctx->stream);
        *out = mrs::build_graph(ctx, idx.p, dist.p, n, k, gp->bandwidth, gp->local_connectivity,
                                gp->mix_weight, n_epochs);
    });
}

int mrs_graph_from_coo(mrs_ctx* ctx, int64_t n, const int32_t* rows, const int32_t* cols,
                       const float* vals, int64_t nnz, int32_t n_epochs, mrs_graph** out) {
    return mrs::guarded([&] {
        if (!ctx || !out || (nnz > 0 && (!rows || !cols || !vals)))
            throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        MRS_HIP(hipSetDevice(ctx->device));
        mrs::StreamScope scope(ctx->stream, ctx->pool);
        *out = mrs::graph_from_coo(ctx, n, rows, cols, vals, nnz, n_epochs);
    });
}

int mrs_graph_destroy(mrs_graph* g) {
    return mrs::guarded([&] {
        if (!g) return;
        (void)hipStreamSynchronize(g->ctx->stream);
        mrs::StreamScope scope(g->ctx->stream, g->ctx->pool);
        delete g;
    });
}

int mrs_graph_info(const mrs_graph* g, int64_t* n, int64_t* nnz, int64_t* n_edges) {
    return mrs::guarded([&] {
        if (!g) throw Error(MRS_INVALID_ARGUMENT, "graph is NULL");
        if (n) *n = g->n;
        if (nnz) *nnz = g->nnz;
        if (n_edges) *n_edges = g->n_edges;
    });
}

int mrs_graph_export_coo(const mrs_graph* g, int32_t* rows, int32_t* cols, float* vals) {
    return mrs::guarded([&] {
        if (!g) throw Error(MRS_INVALID_ARGUMENT, "graph is NULL");
        hipStream_t s = g->ctx->stream;
        const size_t b = sizeof(uint32_t) * (size_t)g->nnz;
        if (g->nnz) {
            if (rows) MRS_HIP(hipMemcpyAsync(rows, g->coo_row.p, b, hipMemcpyDefault, s));
            if (cols) MRS_HIP(hipMemcpyAsync(cols, g->coo_col.p, b, hipMemcpyDefault, s));
            if (vals) MRS_HIP(hipMemcpyAsync(vals, g->coo_val.p, b, hipMemcpyDefault, s));
        }
        MRS_HIP(hipStreamSynchronize(s));
    });
}

int mrs_graph_sigma_rho(const mrs_graph* g, float* sigma, float* rho) {
    return mrs::guarded([&] {
        if (!g) throw Error(MRS_INVALID_ARGUMENT, "graph is NULL");
        if (g->k == 0) throw Error(MRS_INVALID_ARGUMENT, "an imported graph has no sigma/rho");
        hipStream_t s = g->ctx->stream;
        if (sigma) MRS_HIP(hipMemcpyAsync(sigma, g->sigma.p, g->sigma.bytes(), hipMemcpyDefault, s));
        if (rho) MRS_HIP(hipMemcpyAsync(rho, g->rho.p, g->rho.bytes(), hipMemcpyDefault, s));
        MRS_HIP(hipStreamSynchronize(s));
    });
}

int mrs_dens_state_from_graph(mrs_ctx* ctx, const mrs_graph* g, const int32_t* knn_idx,
                              const float* knn_dist, int32_t distances_squared, float* r_out,
                              float* mu_sum_out, float* mu_tot_out) {
    return mrs::guarded([&] {
        if (!ctx || !g || !r_out || !mu_sum_out)
            throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        if (g->k == 0)
            throw Error(MRS_INVALID_ARGUMENT, "densMAP needs a graph built from kNN");
        mrs::StreamScope scope(ctx->stream, ctx->pool);
        const size_t nk = (size_t)g->n * (size_t)g->k;
        mrs::DevView<int32_t> idx(knn_idx, nk, ctx->stream);
        mrs::DevView<float> dist(knn_dist, nk, ctx->stream);
        mrs::DevView<float> r(r_out, (size_t)g->n, ctx->stream, false);
        mrs::DevView<float> ms(mu_sum_out, (size_t)g->n, ctx->stream, false);
        float tot = mrs::dens_state_on_device(const_cast<mrs_graph*>(g), idx.p, dist.p,
                                              distances_squared != 0, r.p, ms.p);
        r.download(ctx->stream);
        ms.download(ctx->stream);
        MRS_HIP(hipStreamSynchronize(ctx->stream));
        if (mu_tot_out) *mu_tot_out = tot;
    });
}

int mrs_umap_state_from_graph(mrs_ctx* ctx, const mrs_graph* g, const float* embd, int32_t n_dim,
                              const mrs_umap_optim_params* params, const mrs_dens_state* dens,
                              mrs_umap_state** out) {
    return mrs::guarded([&] {
        if (!ctx || !g || !embd || !params || !out)
            throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        mrs::StreamScope scope(ctx->stream, ctx->pool);
        if (n_dim <= 0 || n_dim > 16) throw Error(MRS_INVALID_ARGUMENT, "n_dim must be 1..16");
        if (params->n_epochs <= 0) throw Error(MRS_INVALID_ARGUMENT, "n_epochs must be positive");
        if (g->n_edges == 0) throw Error(MRS_NO_GRAPH_EDGES, "UMAP: no edges to optimise (NoGraphEdges)");
        hipStream_t s = ctx->stream;
        auto* st = new mrs_umap_state();
        try {
            st->ctx = ctx;
            st->n = g->n;
            st->dim = n_dim;
            st->n_edges = g->n_edges;
            st->params = *params;
            auto cp = [&](auto& dst, const auto& src, size_t count) {
                dst.alloc(count);
                MRS_HIP(hipMemcpyAsync(dst.p, src.p, dst.bytes(), hipMemcpyDeviceToDevice, s));
            };
            const size_t E = (size_t)g->n_edges;
            cp(st->off, g->off, (size_t)g->n + 1);
            cp(st->csr_edge_idx, g->csr_edge_idx, 2 * E);
            cp(st->csr_other, g->csr_other, 2 * E);
            cp(st->edges, g->edges, E);
            cp(st->per, g->per, E);
            cp(st->next, g->per, E);
            cp(st->mu, g->edge_w, E);
            if (dens) mrs::state_alloc_dens(st, dens->params, dens->r, dens->mu_tot);
            mrs::state_alloc_nodes(st, embd);
            MRS_HIP(hipStreamSynchronize(s));
        } catch (...) {
            delete st;
            throw;
        }
        *out = st;
    });
}

}  // extern "C"
