// Device code of the edge-parallel Adam epoch shared by the single-fit kernels
// (epoch.hip) and the multi-fit kernels of the fit server (serve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "epoch_common.hpp"

namespace mrs {

// x^y for x > 0 (a normal float) from the hardware base-2 logarithm and
// exponential: exp2(y * log2(x)).  Its relative error is about
// ln2 * |y * log2(x)| * 2^-23 plus an ulp, i.e. below 2e-6 for the
// arguments of the epoch and radii kernels (x = a squared distance >= 1e-8,
// or one plus 0.001; y = b in [0.1, 2]), against ~100 instructions for the
// correctly rounded powf.  The node-exact strategy keeps powf.
__device__ inline float pos_powf(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

__device__ inline float clampf(float g, float c) {
    if (g > c) g = c;
    if (g < -c) g = -c;
    return g;
}

// Segmented sum over the wavefront's contiguous runs of equal keys; returns
// whether this lane heads its run, and the head then holds the run's total.
// A run is named by its head lane, not by the key: the same key can recur
// after a gap (inactive lanes in between), and those are separate runs whose
// heads each add their own part -- reducing by key alone would let the
// doubling steps reach across the gap and count the later run twice.
template <int D>
__device__ inline bool seg_reduce(uint32_t key, float (&v)[D]) {
    const int lane = threadIdx.x & 63;
    const uint32_t kprev = __shfl_up(key, 1, 64);
    const bool head = lane == 0 || kprev != key;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    const uint32_t run = 63u - (uint32_t)__clzll(heads & upto);
#pragma unroll
    for (int offd = 1; offd < 64; offd <<= 1) {
        const uint32_t ro = __shfl_down(run, offd, 64);
        const bool take = (lane + offd < 64) && ro == run;
#pragma unroll
        for (int d = 0; d < D; d++) {
            float vo = __shfl_down(v[d], offd, 64);
            if (take) v[d] += vo;
        }
    }
    return head;
}

// The work of one firing undirected edge (i < j) of one Adam epoch: gathers
// y_i, y_j and 2*neg_rate negatives, scatters the larger endpoint's gradient
// with FP32 atomics and the smaller endpoint's after a segmented sum over the
// wavefront (edges arrive sorted by i).  Every lane of the wavefront calls it
// (the reduction is wavefront-wide); lanes with active == false only join the
// reduction.  DIM is the compile-time embedding dimension.
template <int DIM, bool DENS>
__device__ __forceinline__ void edge_work(bool active, uint32_t e, const uint4* __restrict__ edges,
    const float* __restrict__ embd, float* __restrict__ grad, uint32_t* __restrict__ flag,
    uint32_t epoch, uint32_t seed, FastMod fm, float a, float b, float two_a_b,
    float two_gamma_b, int neg_rate, const float* __restrict__ mu,
    const float4* __restrict__ dens_pack, const float* __restrict__ dens_scalars) {
    const float two = 2.0f, one = 1.0f;
    uint32_t key = 0xFFFFFFFFu;
    float gi[DIM], gj[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) gi[d] = gj[d] = 0.0f;
    uint32_t j = 0;
    if (active) {
        const uint4 rec = edges[e];  // {i, j, CSR slot under i, under j}
        const uint32_t i = rec.x;
        j = rec.y;
        key = i;
        const uint32_t li = rec.z, lj = rec.w;
        float yi[DIM], yj[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            yi[d] = embd[(size_t)i * DIM + d];
            yj[d] = embd[(size_t)j * DIM + d];
        }
        float dist_sq = 0.0f;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            float diff = yi[d] - yj[d];
            dist_sq += diff * diff;
        }
        if (dist_sq >= kGradDistSqThreshold) {
            float dsb = pos_powf(dist_sq, b);
            float denom = one + a * dsb;
            float gc = two_a_b * dsb / (dist_sq * denom);
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                float t = two * gc * (yj[d] - yi[d]);
                gi[d] += t;
                gj[d] -= t;
            }
            if (DENS) {
                float phi = one / denom;
                float dphi_term = a * b * dsb / dist_sq * phi;
                float common = one - b * (one - phi);
                // per node (w, 1 / phi_sum, 1 / exp(re)): one 16-byte gather per endpoint
                const float4 pi = dens_pack[i], pj = dens_pack[j];
                float dr_i = (phi * pi.y) * (common * pi.z + dphi_term);
                float dr_j = (phi * pj.y) * (common * pj.z + dphi_term);
                float cor = dens_scalars[3] * (pi.x * dr_i + pj.x * dr_j) / mu[e];
                float two_cor = two * cor;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    float t = two * clampf(two_cor * (yi[d] - yj[d]), kGradClipVal);
                    gi[d] += t;
                    gj[d] -= t;
                }
            }
        }
        // negatives of both endpoints: gather all rows first, then consume
        for (int s = 0; s < neg_rate; s++) {
            uint32_t ki = hash_neg(seed, i, epoch, li, (uint32_t)s, fm);
            uint32_t kj = hash_neg(seed, j, epoch, lj, (uint32_t)s, fm);
            float yki[DIM], ykj[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                yki[d] = embd[(size_t)ki * DIM + d];
                ykj[d] = embd[(size_t)kj * DIM + d];
            }
            if (ki != i) {
                float dsk = 0.0f;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    float diff = yi[d] - yki[d];
                    dsk += diff * diff;
                }
                float safe = dsk + kGradRepEps;
                float gc = clampf(two_gamma_b / (safe * (one + a * pos_powf(safe, b))), kGradClipVal);
#pragma unroll
                for (int d = 0; d < DIM; d++) gi[d] += gc * (yi[d] - yki[d]);
            }
            if (kj != j) {
                float dsk = 0.0f;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    float diff = yj[d] - ykj[d];
                    dsk += diff * diff;
                }
                float safe = dsk + kGradRepEps;
                float gc = clampf(two_gamma_b / (safe * (one + a * pos_powf(safe, b))), kGradClipVal);
#pragma unroll
                for (int d = 0; d < DIM; d++) gj[d] += gc * (yj[d] - ykj[d]);
            }
        }
#pragma unroll
        for (int d = 0; d < DIM; d++) unsafeAtomicAdd(&grad[(size_t)j * DIM + d], gj[d]);
        flag[j] = 1u;
    }
    // smaller endpoint: combine runs of equal i within the wavefront first
    const bool head = seg_reduce<DIM>(key, gi);
    if (head && key != 0xFFFFFFFFu) {
#pragma unroll
        for (int d = 0; d < DIM; d++) unsafeAtomicAdd(&grad[(size_t)key * DIM + d], gi[d]);
        flag[key] = 1u;
    }
}

// One undirected edge of one Adam epoch, for thread-per-edge kernels: e is
// this thread's edge (e >= n_edges: an idle lane that still joins the
// wavefront's reduction).  Reads next/per coalesced and advances a firing
// edge's schedule cursor in place, then does the edge's work.
template <int DIM, bool DENS>
__device__ __forceinline__ void edge_epoch(uint32_t e, const uint4* __restrict__ edges,
    const float* __restrict__ per, float* __restrict__ next, const float* __restrict__ embd,
    float* __restrict__ grad, uint32_t* __restrict__ flag, uint32_t n_edges, uint32_t epoch,
    float epoch_f, uint32_t seed, FastMod fm, float a, float b, float two_a_b,
    float two_gamma_b, int neg_rate, const float* __restrict__ mu,
    const float4* __restrict__ dens_pack, const float* __restrict__ dens_scalars) {
    bool active = false;
    if (e < n_edges) {
        float nx = next[e];
        if (nx <= epoch_f) {
            active = true;
            next[e] = nx + per[e];
        }
    }
    edge_work<DIM, DENS>(active, e, edges, embd, grad, flag, epoch, seed, fm, a, b, two_a_b,
                         two_gamma_b, neg_rate, mu, dens_pack, dens_scalars);
}

// Adam step of one node if flagged (umap_adam_update, umap_optimiser_gpu.rs:887-923),
// then clear its grad and flag for the next accumulate-by-atomics epoch.
template <int DIM>
__device__ __forceinline__ void adam_clear_node(uint32_t node, float* __restrict__ grad,
                                                uint32_t* __restrict__ flag,
                                                float* __restrict__ m, float* __restrict__ v,
                                                float* __restrict__ embd, float fused_lr,
                                                float epsc, float omb1, float omb2,
                                                float* __restrict__ norm_partial) {
    if (flag[node] == 0u) {
        if (norm_partial) norm_partial[node] = 0.0f;
        return;
    }
    float nrm = 0.0f;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const size_t q = (size_t)node * DIM + d;
        float g = grad[q];
        nrm += g * g;
        float mo = m[q];
        float mn = mo + omb1 * (g - mo);
        m[q] = mn;
        float vo = v[q];
        float vn = vo + omb2 * (g * g - vo);
        v[q] = vn;
        embd[q] += fused_lr * mn / (sqrtf(vn) + epsc);
        grad[q] = 0.0f;
    }
    flag[node] = 0u;
    if (norm_partial) norm_partial[node] = nrm;
}

}  // namespace mrs
