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

// grad[node] += v by FP32 atomics for every lane with on == true, all of a
// node's coordinates in the same wave instruction.  FP32 atomics execute
// beyond the L2 (about 21 G requests/s on this GPU whatever the scope or the
// table size), and the atomics of one instruction that fall in one cache line
// leave as one request; a node's coordinates share a line, so issuing them
// together costs one request per node instead of DIM.  Lanes work in groups
// of P >= DIM: in round r, lane L adds coordinate L % P of the node held by
// lane r * (64 / P) + L / P.  Every lane of the wavefront must call this.
template <int DIM>
__device__ inline void atomic_add_rows(float* __restrict__ grad, uint32_t node, bool on,
                                       const float (&v)[DIM]) {
    constexpr int P = DIM == 1 ? 1 : DIM == 2 ? 2 : DIM <= 4 ? 4 : 8;
    static_assert(DIM <= P && P <= 8, "at most 8 coordinates per node");
    constexpr int PER = 64 / P;
    const int lane = threadIdx.x & 63;
    const int c = lane % P;
    const uint64_t busy = __ballot(on);
#pragma unroll
    for (int r = 0; r < P; r++) {
        if (((busy >> (r * PER)) & (PER == 64 ? ~0ull : ((1ull << PER) - 1ull))) == 0ull) continue;
        const int src = r * PER + lane / P;
        const uint32_t nd = __shfl(node, src, 64);
        const bool o = (busy >> src) & 1ull;
        float val = 0.0f;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const float t = __shfl(v[d], src, 64);
            if (d == c) val = t;
        }
        if (o && c < DIM) unsafeAtomicAdd(&grad[(size_t)nd * DIM + c], val);
    }
}

// "No update this epoch" is marked in the accumulated gradient itself: the
// accumulate-by-atomics buffer starts each epoch at -0.0 in every component.
// Every contribution an edge adds is a sum that started at +0.0, and a sum of
// such values is never -0.0 (in round-to-nearest, x + y = -0 only for
// x = y = -0), so one atomic add of any contribution, even an exact zero,
// turns the component into something other than -0.0.  A node is flagged when
// any component differs from -0.0.  (Only a contribution that is a negative
// denormal in every component and flushed by the atomic unit could leave all
// markers standing; contributions are differences of coordinates times
// bounded factors, far from the denormal range.)  This replaces a per-node
// flag array, whose store cost one more scattered request per firing edge.
constexpr uint32_t kNoUpdateBits = 0x80000000u;  // -0.0f

template <int DIM>
__device__ inline bool has_update_mark(const float* g) {
    bool any = false;
#pragma unroll
    for (int d = 0; d < DIM; d++) any |= __float_as_uint(g[d]) != kNoUpdateBits;
    return any;
}

// Repulsive term of one negative sample k of node y (umap_grad_accum's
// negative loop, umap_optimiser_gpu.rs:806-849): skipped for k = the node.
template <int DIM>
__device__ inline void repulse(const float (&y)[DIM], const float (&yk)[DIM], bool self, float a,
                               float b, float two_gamma_b, float (&g)[DIM]) {
    if (self) return;
    const float one = 1.0f;
    float dsk = 0.0f;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        float diff = y[d] - yk[d];
        dsk += diff * diff;
    }
    float safe = dsk + kGradRepEps;
    float gc = clampf(two_gamma_b / (safe * (one + a * pos_powf(safe, b))), kGradClipVal);
#pragma unroll
    for (int d = 0; d < DIM; d++) g[d] += gc * (y[d] - yk[d]);
}

// Negatives of one endpoint (node, CSR slot loc), added to g in s order.
// Their rows are gathered kNegBatch at a time, all in flight together.
constexpr int kNegBatch = 8;
template <int DIM>
__device__ inline void negatives(uint32_t node, uint32_t loc, const float (&y)[DIM],
                                 const float* __restrict__ embd, uint32_t epoch, uint32_t seed,
                                 FastMod fm, float a, float b, float two_gamma_b, int neg_rate,
                                 float (&g)[DIM]) {
    for (int s0 = 0; s0 < neg_rate; s0 += kNegBatch) {
        uint32_t k[kNegBatch];
        float yk[kNegBatch][DIM];
#pragma unroll
        for (int q = 0; q < kNegBatch; q++)
            if (s0 + q < neg_rate) {
                k[q] = hash_neg(seed, node, epoch, loc, (uint32_t)(s0 + q), fm);
#pragma unroll
                for (int d = 0; d < DIM; d++) yk[q][d] = embd[(size_t)k[q] * DIM + d];
            }
#pragma unroll
        for (int q = 0; q < kNegBatch; q++)
            if (s0 + q < neg_rate) repulse<DIM>(y, yk[q], k[q] == node, a, b, two_gamma_b, g);
    }
}

// The attractive (and densMAP) term of edge e between y_i and y_j, added to
// gi and subtracted from gj (umap_grad_accum's edge loop,
// umap_optimiser_gpu.rs:740-800; the negation is exact).
template <int DIM, bool DENS>
__device__ __forceinline__ void attract(uint32_t i, uint32_t j, uint32_t e, const float (&yi)[DIM],
                                        const float (&yj)[DIM], float a, float b, float two_a_b,
                                        const float* __restrict__ mu,
                                        const float4* __restrict__ dens_pack,
                                        const float* __restrict__ dens_scalars, float (&gi)[DIM],
                                        float (&gj)[DIM]) {
    const float two = 2.0f, one = 1.0f;
    float dist_sq = 0.0f;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        float diff = yi[d] - yj[d];
        dist_sq += diff * diff;
    }
    if (dist_sq < kGradDistSqThreshold) return;
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

// The work of one firing undirected edge e (i < j; rec = its record {i, j,
// CSR slot under i, under j}) of one Adam epoch: gathers
// y_i, y_j, computes the attractive (and densMAP) term once for both
// endpoints, then j's negatives, whose sum goes to grad[j] by FP32 atomics
// at once (while y_j's line is still in L2), then i's negatives; i's total
// goes out after a segmented sum over the wavefront (edges arrive sorted by
// i).  Both sums leave through atomic_add_rows, one request per node.  Every
// lane of the wavefront calls it (the reduction and the atomics are
// wavefront-wide); lanes with active == false only join them.  DIM is the
// compile-time embedding dimension.
template <int DIM, bool DENS>
__device__ __forceinline__ void edge_work(bool active, uint32_t e, uint4 rec,
    const float* __restrict__ embd, float* __restrict__ grad, uint32_t epoch, uint32_t seed,
    FastMod fm, float a, float b, float two_a_b, float two_gamma_b, int neg_rate,
    const float* __restrict__ mu, const float4* __restrict__ dens_pack,
    const float* __restrict__ dens_scalars) {
    uint32_t key = 0xFFFFFFFFu, j = 0, slot_i = 0;
    float gi[DIM], gj[DIM], yi[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) gi[d] = gj[d] = 0.0f;
    if (active) {
        const uint32_t i = rec.x;
        j = rec.y;
        key = i;
        slot_i = rec.z;
        float yj[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            yi[d] = embd[(size_t)i * DIM + d];
            yj[d] = embd[(size_t)j * DIM + d];
        }
        attract<DIM, DENS>(i, j, e, yi, yj, a, b, two_a_b, mu, dens_pack, dens_scalars, gi, gj);
        negatives<DIM>(j, rec.w, yj, embd, epoch, seed, fm, a, b, two_gamma_b, neg_rate, gj);
    }
    atomic_add_rows<DIM>(grad, j, active, gj);
    if (active)
        negatives<DIM>(key, slot_i, yi, embd, epoch, seed, fm, a, b, two_gamma_b, neg_rate, gi);
    // smaller endpoint: combine runs of equal i within the wavefront first
    const bool head = seg_reduce<DIM>(key, gi);
    atomic_add_rows<DIM>(grad, key, head && key != 0xFFFFFFFFu, gi);
}

// A wavefront's LDS queue of firing edges: their ids and 16-byte records.
template <int SPAN>
struct SpanQueue {
    uint4 rec[SPAN];
    uint32_t e[SPAN];
};

// One 64-edge round of a queue: the firing lanes (bits of m) append their
// edge and record at count in lane order.  The other lanes write too, into the
// slots just past the firing ones (which the next round's firing lanes
// overwrite, or which lie past the final count and are never read; the last
// round ends at most at SPAN): every lane stores, so the record loads feeding
// the stores are issued unconditionally and all rounds' loads are in flight
// together.
template <int SPAN>
__device__ __forceinline__ void queue_put(SpanQueue<SPAN>& q, uint32_t count, uint64_t m, int lane,
                                          uint32_t e, uint4 rec) {
    const uint64_t below = (1ull << lane) - 1ull;
    const bool fire = (m >> lane) & 1ull;
    const uint32_t slot = count + (fire ? (uint32_t)__popcll(m & below)
                                        : (uint32_t)__popcll(m) + (uint32_t)__popcll(~m & below));
    q.e[slot] = e;
    q.rec[slot] = rec;
}

// The firing edges of one wavefront's span of SPAN consecutive edges
// [base, base + SPAN) of one Adam epoch, queued in edge order in the
// wavefront's LDS queue (ids and records); returns their count
// (wavefront-uniform).  About one edge in five fires in an epoch, scattered
// over the edge list, so with one lane per edge nearly every wavefront would
// have a firing lane and run the whole negative-sampling body for all 64
// lanes; a queue runs it for full batches.  Runs of equal i stay together for
// the segmented sum.  Built from ballots: no counter is shared between
// wavefronts.  Each firing lane loads its edge's record before queueing it,
// so the gathers that follow the queue depend on no further load.  This one
// advances the span's schedule cursors (coalesced; every cursor and period of
// the span is loaded in one round trip).
template <int SPAN>
__device__ __forceinline__ uint32_t queue_from_schedule(uint64_t base, SpanQueue<SPAN>& q,
                                                        const uint4* __restrict__ edges,
                                                        const float* __restrict__ per,
                                                        float* __restrict__ next,
                                                        uint32_t n_edges, float epoch_f) {
    static_assert(SPAN % 64 == 0, "whole wavefront rounds");
    constexpr int R = SPAN / 64;
    const int lane = threadIdx.x & 63;
    float nx[R], pr[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint64_t e = base + (uint64_t)r * 64 + lane;
        nx[r] = INFINITY;  // past the end: never fires
        pr[r] = 0.0f;
        if (e < n_edges) {
            nx[r] = next[e];
            pr[r] = per[e];
        }
    }
    // every lane loads (a lane whose edge does not fire reads the span's first
    // record, which one request serves for all of them): no branch, so the
    // rounds' loads are all in flight together
    uint4 rec[R];
#pragma unroll
    for (int r = 0; r < R; r++)
        rec[r] = edges[nx[r] <= epoch_f ? base + (uint64_t)r * 64 + lane : base];
    uint32_t count = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint64_t e = base + (uint64_t)r * 64 + lane;
        const bool fire = nx[r] <= epoch_f;
        if (fire) next[e] = nx[r] + pr[r];
        const uint64_t m = __ballot(fire);
        queue_put(q, count, m, lane, (uint32_t)e, rec[r]);
        count += (uint32_t)__popcll(m);
    }
    return count;
}

// The same queue from the epoch's plane of a firing bitmask (bit e mod 64 of
// word e / 64 set when edge e fires; `words` words, zero past the last edge),
// written ahead for a window of epochs by an exact replay of the schedule: one
// 8-byte word per 64 edges instead of their cursors and periods.  The records
// come in compact form (CompactRecords), 8 bytes per firing edge instead of
// 16.  base is wavefront-uniform, so the per-word data comes through the
// scalar cache.
struct CompactRecords {
    const uint2* pair;         // [E] {j, CSR slot of the edge under j}
    const uint64_t* run_start; // [words] bit b: edge 64w + b starts a run of equal i
    const uint32_t* word_run;  // [words] (runs started before edge 64w) - 1, mod 2^32
    const uint2* run_info;     // [runs] {i, CSR slot under i of the run's first edge - its index}
};
template <int SPAN>
__device__ __forceinline__ uint32_t queue_from_bits(uint64_t base, SpanQueue<SPAN>& q,
                                                    const CompactRecords cr,
                                                    const uint64_t* __restrict__ plane,
                                                    uint64_t words) {
    static_assert(SPAN % 64 == 0, "whole wavefront rounds");
    constexpr int R = SPAN / 64;
    const int lane = threadIdx.x & 63;
    const uint64_t w0 = base / 64;
    const uint64_t upto = (2ull << lane) - 1ull;  // bits 0..lane (all 64 for lane 63)
    uint64_t m[R], rs[R];
    uint32_t wr[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const bool in = w0 + r < words;
        m[r] = in ? plane[w0 + r] : 0ull;
        rs[r] = in ? cr.run_start[w0 + r] : 0ull;
        wr[r] = in ? cr.word_run[w0 + r] : 0u;
    }
    // Every lane loads (a lane whose edge does not fire reads the record of
    // the span's first edge, which always exists and which one request serves
    // for all such lanes): no branch, so the rounds' loads are all in flight
    // together.  Edges are sorted by i and a run's slots under i rise by one
    // per edge, so an edge's run is word_run + its marks up to its bit, and
    // its slot under i is its index plus the run's offset.
    uint2 pj[R], ri[R];
    const uint32_t run0 = wr[0] + (uint32_t)(rs[0] & 1ull);  // the run of edge base
#pragma unroll
    for (int r = 0; r < R; r++) {
        const bool fire = (m[r] >> lane) & 1ull;
        pj[r] = cr.pair[fire ? base + (uint64_t)r * 64 + lane : base];
        ri[r] = cr.run_info[fire ? wr[r] + (uint32_t)__popcll(rs[r] & upto) : run0];
    }
    uint32_t count = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t e = (uint32_t)(base + (uint64_t)r * 64 + lane);
        queue_put(q, count, m[r], lane, e, make_uint4(ri[r].x, pj[r].x, e + ri[r].y, pj[r].y));
        count += (uint32_t)__popcll(m[r]);
    }
    return count;
}

// Works through a wavefront's queue of `count` firing edges 64 at a time
// (edge_work).  Every wavefront of the workgroup must call this: it holds a
// workgroup barrier between building and reading the queues.
template <int DIM, bool DENS, int SPAN>
__device__ __forceinline__ void work_queue(const SpanQueue<SPAN>& q, uint32_t count,
    const float* __restrict__ embd, float* __restrict__ grad,
    uint32_t epoch, uint32_t seed, FastMod fm, float a, float b, float two_a_b,
    float two_gamma_b, int neg_rate, const float* __restrict__ mu,
    const float4* __restrict__ dens_pack, const float* __restrict__ dens_scalars) {
    const int lane = threadIdx.x & 63;
    __syncthreads();  // the queue is read by other lanes than the ones that wrote it
    for (uint32_t q0 = 0; q0 < count; q0 += 64) {
        const uint32_t slot = q0 + lane;
        const bool active = slot < count;
        const uint32_t e = active ? q.e[slot] : 0u;
        const uint4 rec = active ? q.rec[slot] : make_uint4(0u, 0u, 0u, 0u);
        edge_work<DIM, DENS>(active, e, rec, embd, grad, epoch, seed, fm, a, b, two_a_b,
                             two_gamma_b, neg_rate, mu, dens_pack, dens_scalars);
    }
}

// One wavefront's span of one Adam epoch: queue from the schedule cursors,
// then work through it.  Every wavefront of the workgroup must call this.
template <int DIM, bool DENS, int SPAN>
__device__ __forceinline__ void epoch_edge_span(uint64_t base, SpanQueue<SPAN>& q,
    const uint4* __restrict__ edges, const float* __restrict__ per, float* __restrict__ next,
    const float* __restrict__ embd, float* __restrict__ grad, uint32_t n_edges, uint32_t epoch,
    float epoch_f, uint32_t seed, FastMod fm, float a, float b, float two_a_b,
    float two_gamma_b, int neg_rate, const float* __restrict__ mu,
    const float4* __restrict__ dens_pack, const float* __restrict__ dens_scalars) {
    const uint32_t count = queue_from_schedule<SPAN>(base, q, edges, per, next, n_edges, epoch_f);
    work_queue<DIM, DENS, SPAN>(q, count, embd, grad, epoch, seed, fm, a, b, two_a_b,
                                two_gamma_b, neg_rate, mu, dens_pack, dens_scalars);
}

// Adam step of one node if it got an update this epoch (umap_adam_update,
// umap_optimiser_gpu.rs:887-923), then its gradient back to the -0.0 marker
// for the next accumulate-by-atomics epoch.
template <int DIM>
__device__ __forceinline__ void adam_clear_node(uint32_t node, float* __restrict__ grad,
                                                float* __restrict__ m, float* __restrict__ v,
                                                float* __restrict__ embd, float fused_lr,
                                                float epsc, float omb1, float omb2,
                                                float* __restrict__ norm_partial) {
    float g[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) g[d] = grad[(size_t)node * DIM + d];
    if (!has_update_mark<DIM>(g)) {
        if (norm_partial) norm_partial[node] = 0.0f;
        return;
    }
    float nrm = 0.0f;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const size_t q = (size_t)node * DIM + d;
        nrm += g[d] * g[d];
        float mo = m[q];
        float mn = mo + omb1 * (g[d] - mo);
        m[q] = mn;
        float vo = v[q];
        float vn = vo + omb2 * (g[d] * g[d] - vo);
        v[q] = vn;
        embd[q] += fused_lr * mn / (sqrtf(vn) + epsc);
        grad[q] = -0.0f;
    }
    if (norm_partial) norm_partial[node] = nrm;
}

}  // namespace mrs
