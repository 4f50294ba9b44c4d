// Exhaustive kNN for the "exhaustive_gpu" path (src/data/nearest_neighbours_gpu.rs:282-290,
// self filter and take-k :350-362): squared Euclidean, self excluded, ascending,
// ties broken by index.
//
// Stage 1 -- candidate shortlist by a distance GEMM on MFMA (gfx950):
//   score(q, j) = x_q . y_j - 0.5 |y_j|^2      (= -0.5 |x_q - y_j|^2 + const(q))
//   The accumulator is initialised with -0.5 |y_j|^2, so one v_mfma chain gives
//   the score and ranking needs only a max/compare per element.  The database
//   is the MFMA A operand (32 rows per tile) and the queries the B operand, so a
//   lane's 16 accumulator values all belong to ONE query (column lane%32): the
//   per-query threshold is a single register and the common case (no element
//   beats it) costs 15 maxes + 1 compare + 1 ballot per 16 candidates.
//   Precision FP32 (v_mfma_f32_32x32x2_f32) or BF16 (v_mfma_f32_32x32x16_bf16,
//   FP32 accumulate).
//   Each workgroup holds 4 waves x 32 queries; database stages of 32-128 rows
//   are staged through two LDS buffers (rows padded by 16 B against bank
//   conflicts), prefetched one stage ahead in registers and shared by the 4
//   waves.  Candidates beating the query's threshold go to a per-query queue
//   (64 entries, in global memory: it is touched only on the rare path); a
//   queue that could overflow on the next tile is merged into the query's
//   sorted top-S list by the whole wavefront (bitonic sort + merge on 64-bit
//   keys), which also refreshes the threshold.  The query's own row is
//   dropped at insertion.  Problems below 65 536 rows split the database so
//   the grid still fills 256 CUs, and a merge kernel joins the split lists.
//
// Pruned sweep (65 536 rows and more; the result is the same, bit for bit):
//   rows are put in cell order (nearest of n/1024 seed rows, stable radix sort
//   by cell, cells padded to whole query blocks), each query block first
//   sweeps its own cell, and then only the stages whose centroid/radius bound
//   cannot rule them out for any of its queries (see the pruned-sweep kernels
//   below).  The re-rank then runs in cell order as well.
//
// Stage 2 -- exact re-rank: one wavefront per query recomputes the S candidate
//   distances in FP32 as fmaf((x_q[t]-x_c[t])^2) over t in order (the same
//   arithmetic as the CPU oracle) and sorts (distance, index) keys; the first k
//   are the result.  This makes the indices exact whenever the true top-k is in
//   the shortlist, for both precisions.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include <algorithm>
#include <numeric>
#include <vector>
#include <type_traits>
#include <cfloat>
#include <cmath>

#include "autotune.hpp"
#include "knn.hpp"

namespace mrs {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef unsigned short u16;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;  // 16-byte staging chunk

constexpr int kWavesMax = 8;       // most waves per workgroup in the shortlist kernel
constexpr int kQPerWave = 32;      // queries per wave (MFMA B columns)
constexpr int kSplitAlign = 128;   // database splits are multiples of every stage size
constexpr int kStageChunks = 2;    // 16-byte chunks per thread at kWavesMax waves (1024 per stage)
constexpr uint64_t kInfKey = ~0ull;

// ---------------------------------------------------------------- keys

// Monotone map float -> u32 (larger float -> larger u32).
__device__ inline uint32_t ord_f32(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float unord_f32(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}
// ascending key order == descending score, then ascending index
__device__ inline uint64_t score_key(float score, uint32_t idx) {
    return ((uint64_t)(~ord_f32(score)) << 32) | idx;
}
__device__ inline float key_score(uint64_t key) {
    return key == kInfKey ? -INFINITY : unord_f32(~(uint32_t)(key >> 32));
}

// v from lane (lane ^ m).  The sorting networks below only ask for constant
// masks (their loops are fully unrolled), and each mask takes the cheapest
// cross-lane path: DPP quad permutes and row mirrors (VALU) for 1, 2, 3, 7, 15;
// ds_swizzle's xor mode (32-lane groups, no LDS memory access) for 4, 8, 16, 31;
// ds_bpermute only for the masks that cross the two 32-lane halves.
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, int m) {
    switch (m) {
        case 1: return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);   // quad_perm [1,0,3,2]
        case 2: return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);   // quad_perm [2,3,0,1]
        case 3: return __builtin_amdgcn_mov_dpp(v, 0x1B, 0xF, 0xF, true);   // quad_perm [3,2,1,0]
        case 7: return __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true);  // row_half_mirror
        case 15: return __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true); // row_mirror
        case 4: return __builtin_amdgcn_ds_swizzle(v, 0x1F | (4 << 10));
        case 8: return __builtin_amdgcn_ds_swizzle(v, 0x1F | (8 << 10));
        case 16: return __builtin_amdgcn_ds_swizzle(v, 0x1F | (16 << 10));
        case 31: return __builtin_amdgcn_ds_swizzle(v, 0x1F | (31 << 10));
        default: return __shfl_xor(v, m, 64);
    }
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    const uint32_t lo = lane_xor((uint32_t)v, m), hi = lane_xor((uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}

// Sort E*64 keys held as v[e] at global position g = e*64 + lane, ascending,
// with the "flip" bitonic network (minimum always to the lower position).
template <int E>
__device__ inline void wave_sort(uint64_t (&v)[E]) {
    const int lane = threadIdx.x & 63;
    constexpr int N = 64 * E;
#pragma unroll
    for (int s = 2; s <= N; s <<= 1) {
        // flip step: partner g ^ (s-1)
        {
            uint64_t nv[E];
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int m = s - 1;
                const int pe = e ^ (m >> 6);
                const uint64_t o = shfl_xor64(v[pe], m & 63);
                const int g = e * 64 + lane;
                const bool lower = (g & (s - 1)) < (s >> 1);
                nv[e] = lower ? (o < v[e] ? o : v[e]) : (o > v[e] ? o : v[e]);
            }
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = nv[e];
        }
#pragma unroll
        for (int j = s >> 2; j >= 1; j >>= 1) {
            uint64_t nv[E];
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int pe = e ^ (j >> 6);
                const uint64_t o = (j >= 64) ? v[pe] : shfl_xor64(v[e], j);
                const int g = e * 64 + lane;
                const bool lower = (g & j) == 0;
                nv[e] = lower ? (o < v[e] ? o : v[e]) : (o > v[e] ? o : v[e]);
            }
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = nv[e];
        }
    }
}

// Merge a sorted 64-key queue q (one per lane) into a sorted list of 64*E keys
// (v[e] at e*64+lane), keeping the 64*E smallest, sorted.
template <int E>
__device__ inline void wave_merge(uint64_t (&v)[E], uint64_t q) {
    const int lane = threadIdx.x & 63;
    // reversed queue against the top of the list: min() leaves a bitonic sequence
    const uint32_t qh = __shfl((uint32_t)(q >> 32), 63 - lane, 64);
    const uint32_t ql = __shfl((uint32_t)q, 63 - lane, 64);
    const uint64_t qr = ((uint64_t)qh << 32) | ql;
    v[E - 1] = qr < v[E - 1] ? qr : v[E - 1];
    constexpr int N = 64 * E;
#pragma unroll
    for (int j = N >> 1; j >= 1; j >>= 1) {
        uint64_t nv[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int pe = e ^ (j >> 6);
            const uint64_t o = (j >= 64) ? v[pe] : shfl_xor64(v[e], j);
            const int g = e * 64 + lane;
            const bool lower = (g & j) == 0;
            nv[e] = lower ? (o < v[e] ? o : v[e]) : (o > v[e] ? o : v[e]);
        }
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = nv[e];
    }
}

// ------------------------------------------------------------ packing

__device__ inline u16 f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__device__ inline float bf16_to_f32(u16 h) { return __uint_as_float((uint32_t)h << 16); }

// X [n][d] f32 -> P [npad][KP] (bf16 or f32, zero padded) and
// negh[j] = -0.5 |p_j|^2 (-inf for padding rows), computed from the packed values.
template <bool BF16>
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ x, uint32_t n, int d,
                                              uint32_t npad, int KP, void* __restrict__ out,
                                              float* __restrict__ negh) {
    const uint32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= npad) return;
    float part = 0.0f;
    for (int t = lane; t < KP; t += 64) {
        float v = (row < n && t < d) ? x[(size_t)row * d + t] : 0.0f;
        if (BF16) {
            u16 h = f32_to_bf16_rne(v);
            reinterpret_cast<u16*>(out)[(size_t)row * KP + t] = h;
            v = bf16_to_f32(h);
        } else {
            reinterpret_cast<float*>(out)[(size_t)row * KP + t] = v;
        }
        part += v * v;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) part += __shfl_xor(part, o, 64);
    if (lane == 0) negh[row] = row < n ? -0.5f * part : -INFINITY;
}

// ------------------------------------------------------ shortlist kernel

struct ShortlistArgs {
    const void* P;        // packed [npad][KP]
    const float* negh;    // [npad]
    uint64_t* lists;      // [splits][npad][S]
    uint64_t* queue;      // [splits][npad][QC] per-query candidate queues
    uint32_t n, npad;
    int KP;
    uint32_t db_per_split;  // multiple of kSplitAlign
    // Pruned sweep (one split): rows are in cell order, rowmap[p] is the
    // caller's row at position p (list keys carry caller rows), and workgroup
    // g does work item items[g] = (query block b, list offset o, count c, aux
    // slot a): it sweeps the c stages stage_ids[o + t] for block b.  With
    // a == kMainList it keeps the block's own lists (continuing them when
    // cont); otherwise it fills the aux lists of slot a from empty, above the
    // floor thresholds (the block's thresholds when the pass started), and
    // k_merge_aux later folds them into the block's lists.
    const uint32_t* rowmap;
    const uint4* items;
    const uint32_t* stage_ids;
    int cont;
    uint64_t* aux_lists;   // [slots][128][S]
    uint64_t* aux_queue;   // [slots][128][QC]
    const float* floor;    // [npad] or null
};
constexpr uint32_t kMainList = 0xFFFFFFFFu;

// Wave-wide merges of queries' queues into their sorted global lists.  A
// query's queue is two half-queues of QC/2 entries, one per accumulator half
// (lanes b and b + 32), holding c0 and c1 entries: lane l reads entry l % 32
// of half l / 32; the wave sorts the queue and merges it into the list, whose
// new last entry is the query's threshold key.  With a rowmap, queue entries
// carry database positions, which become the caller's rows here (one gather
// per merge rather than one dependent load per queued candidate), so the list
// keys, and the (score, row) order of the merge, are the caller's.
//   One call merges up to kMergeBatch queries bs[0..nb): every queue, list
// and rowmap load of the batch is issued before the first sort, so a batch
// waits out two memory latencies rather than two per query, and the sorts
// then run back to back.  qn holds the lanes' half-queue counts (lanes b and
// b + 32 hold query b's); tk[j] returns query bs[j]'s new threshold key.
constexpr int kMergeBatch = 4;
template <int E, int QC>
__device__ inline void merge_queries(const int (&bs)[kMergeBatch], int nb, uint32_t qn, int lane,
                                     uint64_t* __restrict__ lists,
                                     const uint64_t* __restrict__ myq,
                                     const uint32_t* __restrict__ rowmap,
                                     uint64_t (&tk)[kMergeBatch]) {
    static_assert(QC == 64, "two half-queues of 32: one entry per lane");
    constexpr int S = 64 * E;
    // the queue entries were stored by this wave's lanes: complete those
    // stores before any lane reads another lane's entry
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    uint64_t one[kMergeBatch][1], v[kMergeBatch][E];
#pragma unroll
    for (int j = 0; j < kMergeBatch; j++) {
        one[j][0] = kInfKey;
        if (j < nb) {  // wave-uniform
            const int b = bs[j];
            const uint32_t c = lane < 32 ? __builtin_amdgcn_readlane(qn, b)
                                         : __builtin_amdgcn_readlane(qn, b + 32);
            if ((uint32_t)(lane & 31) < c) one[j][0] = myq[b * QC + lane];
#pragma unroll
            for (int e = 0; e < E; e++) v[j][e] = lists[(size_t)b * S + e * 64 + lane];
        }
    }
    if (rowmap) {
#pragma unroll
        for (int j = 0; j < kMergeBatch; j++)
            if (j < nb && one[j][0] != kInfKey)
                one[j][0] = (one[j][0] & 0xFFFFFFFF00000000ull) | rowmap[(uint32_t)one[j][0]];
    }
    __builtin_amdgcn_sched_barrier(0);  // every load of the batch ahead of the sorts
#pragma unroll
    for (int j = 0; j < kMergeBatch; j++) {
        if (j >= nb) break;  // wave-uniform
        const int b = bs[j];
        wave_sort<1>(one[j]);
        wave_merge<E>(v[j], one[j][0]);
#pragma unroll
        for (int e = 0; e < E; e++) lists[(size_t)b * S + e * 64 + lane] = v[j][e];
        const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v[j][E - 1] >> 32), 63);
        const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v[j][E - 1], 63);
        tk[j] = ((uint64_t)hi << 32) | lo;
    }
}

// Pops up to kMergeBatch query indices (lowest first) from the bit set need.
__device__ inline int pop_queries(uint64_t& need, int (&bs)[kMergeBatch]) {
    int nb = 0;
#pragma unroll
    for (int j = 0; j < kMergeBatch; j++) {
        bs[j] = 0;
        if (need) {
            bs[j] = __builtin_ctzll(need);
            need &= need - 1;
            nb = j + 1;
        }
    }
    return nb;
}

// Epilogue of one 32-row MFMA tile: the lane's 16 scores (one query, rows
// rowbase + (r/4)*8 + h*4 + r%4) against the query's threshold thr (the score
// of its list's last entry, never below floor).  Each of a query's two lanes
// (h = 0, 1) appends to its own half-queue with its own count qn (a register),
// so an insertion needs no count of the tile's candidates and no exchange
// between the lanes; the threshold is identical in both.  Most tiles leave on
// the max/ballot test.  A query with a half-queue that could overflow on the
// next tile (more than QC/2 - 16 entries) is merged.
//   Row-order sweep (PR = false): the stages come in row order, so a score
// equal to the threshold can only belong to a larger row and is not queued.
// Cell-order sweep (PR = true): an equal score is queued too and the merge,
// which keeps the S smallest (score, row) keys, decides the tie by row; queued
// keys carry the database position, which the merge maps to the caller's row
// (rowmap).  In the tile that holds the wave's own query rows (self_tile, a
// wave-uniform flag) the query's own score is set to -inf before any score is
// queued; it is the largest score of the tile, so that tile always gets past
// the max/ballot test, and the test costs nothing extra elsewhere.  In the
// pruned sweep the threshold never drops below -FLT_MAX, so padding rows
// (-inf) fail the test like every row below the threshold: one compare per
// score decides.  gm and m (group_max) come in from the caller, which computes
// them while the next tile's MFMAs run.
__device__ inline void mask_self(f32x16& acc, int h, int col);

__device__ inline void group_max(const f32x16& acc, float (&gm)[4], float& m) {
    // maxima of the four 4-row groups (rows i*8 + h*4 + 0..3), then of the tile
#pragma unroll
    for (int g = 0; g < 4; g++)
        gm[g] = fmaxf(fmaxf(acc[g * 4], acc[g * 4 + 1]), fmaxf(acc[g * 4 + 2], acc[g * 4 + 3]));
    m = fmaxf(fmaxf(gm[0], gm[1]), fmaxf(gm[2], gm[3]));
}

template <int E, int QC, bool PR>
__device__ inline void epilogue(f32x16 acc, const float (&gm)[4], float m, bool self_tile,
                                uint32_t rowbase, int h, int col, int lane, float& thr,
                                uint32_t& qn, uint64_t* __restrict__ lists,
                                uint64_t* __restrict__ myq, const uint32_t* __restrict__ rowmap,
                                float floor) {
    auto reach = [&](float v) { return PR ? v >= thr : v > thr; };
    if (__ballot(reach(m)) == 0ull) return;
    if (self_tile) mask_self(acc, h, col);
    auto take = [&](int r) { return reach(acc[r]); };
    uint64_t* dst = myq + col * QC + h * (QC / 2);
    // a group no element of which reaches the threshold is skipped as a whole
#pragma unroll
    for (int g = 0; g < 4; g++)
        if (reach(gm[g])) {
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (take(g * 4 + j)) {
                    const uint32_t row = rowbase + g * 8 + h * 4 + j;
                    dst[qn++] = score_key(acc[g * 4 + j], row);
                }
        }
    const uint64_t full = __ballot(qn > (uint32_t)(QC / 2 - 16));
    uint64_t need = (full | (full >> 32)) & 0xFFFFFFFFull;  // one bit per query
    while (need) {
        int bs[kMergeBatch];
        const int nb = pop_queries(need, bs);
        uint64_t tk[kMergeBatch];
        merge_queries<E, QC>(bs, nb, qn, lane, lists, myq, PR ? rowmap : nullptr, tk);
#pragma unroll
        for (int j = 0; j < kMergeBatch; j++)
            if (j < nb && col == bs[j]) {
                qn = 0;
                thr = fmaxf(key_score(tk[j]), floor);  // an aux list never goes below its floor
            }
    }
}

// Score of the lane's query row to -inf in the tile that holds the wave's
// query rows: element r is row (r/4)*8 + h*4 + r%4 of the tile, the query's
// row is col.
__device__ inline void mask_self(f32x16& acc, int h, int col) {
#pragma unroll
    for (int r = 0; r < 16; r++)
        if ((r >> 2) * 8 + h * 4 + (r & 3) == col) acc[r] = -INFINITY;
}

// One RT-row stage copied from global memory straight into its LDS buffer
// (global_load_lds: each wave instruction writes 64 consecutive 16-byte LDS
// slots, lane i at slot i), with no registers in between.  The buffer keeps
// rows padded to row_bytes against bank conflicts, so slot p of the stage is
// chunk p % spr of row p / spr (spr = row_bytes / 16 slots per row); the
// padding slot of a row receives a second copy of the row's chunk 0 and is
// never read.  goff[j] is this lane's source offset within the stage for its
// copy j (slot wave*64 + j*W*64 + lane), -1 past the stage.  The stage's RT
// values of -0.5|y|^2 go to tile_negh the same way, 4 bytes per lane.
#define MRS_LDS(p) ((__attribute__((address_space(3))) void*)(p))
template <int NCOPY, int RT, int W>
__device__ inline void stage_copy(const unsigned char* __restrict__ P, const float* __restrict__ negh,
                                  uint32_t row0, int row_elems_bytes, const int* goff,
                                  unsigned char* tile, float* tile_negh, int wave, int lane) {
    const unsigned char* src = P + (size_t)row0 * row_elems_bytes;
#pragma unroll
    for (int j = 0; j < NCOPY; j++) {
        const int base = wave * 64 + j * W * 64;  // wave-uniform
        if (goff[j] >= 0) __builtin_amdgcn_global_load_lds(src + goff[j], MRS_LDS(tile + base * 16), 16, 0, 0);
    }
    for (int base = wave * 64; base < RT; base += W * 64)  // wave-uniform
        if (base + lane < RT)
            __builtin_amdgcn_global_load_lds(negh + row0 + base + lane, MRS_LDS(tile_negh + base), 4, 0, 0);
}

// The distance GEMM + top-S filter.  One workgroup = W waves x 32 queries; the
// database split [db0, db1) (PR: the work item's list of stages, in cell
// order) streams through two LDS stage buffers of RT rows (double buffering:
// one barrier per stage), each stage copied straight into LDS one stage ahead
// (global_load_lds).  QC = per-query queue capacity.  STEPS > 0
// compiles the K-step count in (BF16: KP / 16), so the query fragments take
// only the registers they need and the MFMA chain is straight-line code.
//   The loop is latency-bound (waves mostly wait on LDS, the stage barrier
// and the prefetch), so occupancy pays: the BF16 kernels with the step count
// compiled in are held to 128 registers (four waves per SIMD; at most a few
// spilled dwords, none for the 128-row stages), the FP32 ones to 168 (three).
template <bool BF16, int E, int RT, int QC, int W, int STEPS, bool PR>
__global__ __launch_bounds__(W * 64)
__attribute__((amdgpu_waves_per_eu(BF16 && STEPS > 0 ? 4 : (BF16 ? 1 : 3)))) void k_shortlist(ShortlistArgs a) {
    // copies per thread and stage: the stage's 16-byte chunks (at most
    // kStageChunks * kWavesMax * 64) plus one padding slot per row
    constexpr int NCOPY = (kStageChunks * kWavesMax * 64 + RT + W * 64 - 1) / (W * 64);
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int S = 64 * E;
    constexpr int elem = BF16 ? 2 : 4;
    const int KP = a.KP;
    const int row_bytes = KP * elem + 16;
    const int buf_bytes = RT * row_bytes;
    unsigned char* tile0 = smem;
    unsigned char* tile1 = smem + buf_bytes;
    float* negh0 = reinterpret_cast<float*>(smem + 2 * buf_bytes);
    float* negh1 = negh0 + RT;

    const int tid = threadIdx.x;
    const int lane = tid & 63, h = lane >> 5, col = lane & 31;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    uint4 item = make_uint4(blockIdx.x, 0u, 0u, kMainList);
    if (PR) {
        // the same item for every lane: readfirstlane tells the compiler, so
        // the stage loop, its bounds and the list addresses are scalar
        const uint4 it = a.items[blockIdx.x];
        item = make_uint4(__builtin_amdgcn_readfirstlane(it.x), __builtin_amdgcn_readfirstlane(it.y),
                          __builtin_amdgcn_readfirstlane(it.z), __builtin_amdgcn_readfirstlane(it.w));
    }
    const uint32_t blk_id = item.x;
    const bool aux = PR && item.w != kMainList;
    const uint32_t qbase = (blk_id * W + wave) * kQPerWave;
    const uint32_t split = blockIdx.y;
    const uint32_t db0 = split * a.db_per_split;
    const uint32_t db1 = min(db0 + a.db_per_split, a.npad);
    uint64_t* lists = a.lists + ((size_t)split * a.npad + qbase) * S;
    uint64_t* myq = a.queue + ((size_t)split * a.npad + qbase) * QC;
    if (aux) {
        const size_t r0 = (size_t)item.w * (W * kQPerWave) + (size_t)wave * kQPerWave;
        lists = a.aux_lists + r0 * S;
        myq = a.aux_queue + r0 * QC;
    }

    if (!PR || !a.cont || aux)
        for (int i = lane; i < kQPerWave * S; i += 64) lists[i] = kInfKey;
    // B fragments: this lane's query (column) over its half of K
    const uint32_t q = qbase + col;
    constexpr int MAXSTEPS = STEPS > 0 ? STEPS : (BF16 ? 8 : 64);  // KP <= 128
    const int steps = STEPS > 0 ? STEPS : (BF16 ? KP / 16 : KP / 2);
    bf16x8 bq[BF16 ? MAXSTEPS : 1];
    float bqf[BF16 ? 1 : MAXSTEPS];
    if (BF16) {
        const bf16x8* src = reinterpret_cast<const bf16x8*>(
            reinterpret_cast<const u16*>(a.P) + (size_t)q * KP + h * (KP / 2));
#pragma unroll
        for (int s = 0; s < MAXSTEPS; s++)
            if (s < steps) bq[s] = src[s];
    } else {
        const float* src = reinterpret_cast<const float*>(a.P) + (size_t)q * KP + h * (KP / 2);
#pragma unroll
        for (int s = 0; s < MAXSTEPS; s++)
            if (s < steps) bqf[s] = src[s];
    }
    // a pass that continues: the thresholds left by the earlier pass (an aux
    // item starts from the floor, its lists being empty)
    // pruned sweep: thresholds never below -FLT_MAX (see epilogue)
    const float floor = PR ? fmaxf(a.floor ? a.floor[q] : -INFINITY, -FLT_MAX) : -INFINITY;
    float thr = floor;
    if (PR && a.cont && !aux) thr = fmaxf(key_score(lists[(size_t)col * S + S - 1]), floor);
    uint32_t qn = 0;

    const unsigned char* P = reinterpret_cast<const unsigned char*>(a.P);
    const int row_elems_bytes = KP * elem;
    const int cpr = row_elems_bytes / 16;  // 16-byte chunks per row
    const int spr = cpr + 1;               // LDS slots per padded row
    int goff[NCOPY];
#pragma unroll
    for (int j = 0; j < NCOPY; j++) {
        const int slot = wave * 64 + j * W * 64 + lane;
        const int row = slot / spr, c = slot - row * spr;
        goff[j] = slot < RT * spr ? row * row_elems_bytes + (c < cpr ? c : 0) * 16 : -1;
    }
    // the stages this workgroup sweeps: the split's consecutive stages, or
    // (pruned sweep) its list
    uint32_t nst = (db1 - db0) / RT;
    const uint32_t* sid = nullptr;
    if (PR) {
        nst = item.z;
        sid = a.stage_ids + item.y;
    }
    // The list's stage ids come 64 at a time in a register (lane i: stage
    // c*64 + i of chunk c; idn holds chunk c + 1, loaded a chunk ahead), and
    // readlane hands each id to the scalar unit: no stage waits on a load of
    // its id, and no wait on such a load also waits on the stage copy in flight.
    auto load_ids = [&](uint32_t t0) -> uint32_t { return t0 + lane < nst ? sid[t0 + lane] : 0u; };
    uint32_t idc = 0, idn = 0;
    if (PR) {
        idc = load_ids(0);
        idn = load_ids(64);
    }
    // row of stage t, for t in the current chunk or the next one
    auto srow = [&](uint32_t t, bool next_chunk) -> uint32_t {
        if (!PR) return db0 + t * RT;
        const uint32_t v = __builtin_amdgcn_readlane(next_chunk ? idn : idc, (int)(t & 63u));
        return v * (uint32_t)RT;
    };
    if (nst > 0) stage_copy<NCOPY, RT, W>(P, a.negh, srow(0, false), row_elems_bytes, goff, tile0, negh0, wave, lane);
    // everything loaded so far (query fragments, stage 0) is complete here
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    // One stage: stage st+1 goes straight into the other buffer (last read in
    // stage st-1, before the barrier that closed it) and lands while stage st
    // computes.  The buffers come in as distinct restrict pointers so that the
    // compiler knows the LDS reads of this stage cannot see the copy in flight
    // and does not wait for it before reading.
    auto step = [&](const unsigned char* __restrict__ cur, const float* __restrict__ curn,
                    unsigned char* __restrict__ nxt, float* __restrict__ nxtn, uint32_t st) {
        if (st + 1 < nst)
            stage_copy<NCOPY, RT, W>(P, a.negh, srow(st + 1, ((st + 1) & 63u) == 0u),
                                     row_elems_bytes, goff, nxt, nxtn, wave, lane);
        const uint32_t r0 = srow(st, false);
        if constexpr (BF16 && STEPS > 0 && RT >= 64) {
            // Tiles in a software pipeline: tile t + 1's bias loads and MFMA
            // chain are issued before tile t's group maxima, so the maxima's
            // VALU work runs while the MFMAs do, in this wave as in others.
            constexpr int NT = RT / 32;
            f32x16 acc[2];
            auto init_tile = [&](f32x16& c, int t) {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float4 nv = *reinterpret_cast<const float4*>(curn + t * 32 + i * 8 + h * 4);
                    c[i * 4 + 0] = nv.x;
                    c[i * 4 + 1] = nv.y;
                    c[i * 4 + 2] = nv.z;
                    c[i * 4 + 3] = nv.w;
                }
            };
            auto chain_tile = [&](f32x16& c, int t) {
                const bf16x8* ap = reinterpret_cast<const bf16x8*>(cur + (t * 32 + col) * row_bytes + h * KP);
#pragma unroll
                for (int s = 0; s < STEPS; s++)
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[s], bq[s], c, 0, 0, 0);
            };
            init_tile(acc[0], 0);
            chain_tile(acc[0], 0);
#pragma unroll
            for (int t = 0; t < NT; t++) {
                if (t + 1 < NT) {
                    init_tile(acc[(t + 1) & 1], t + 1);
                    chain_tile(acc[(t + 1) & 1], t + 1);
                }
                float gm[4], m;
                group_max(acc[t & 1], gm, m);
                epilogue<E, QC, PR>(acc[t & 1], gm, m, r0 + t * 32 == qbase, r0 + t * 32, h, col,
                                    lane, thr, qn, lists, myq, a.rowmap, floor);
            }
        } else {
#pragma unroll
            for (int sp = 0; sp < RT / 32; sp += 2) {
                // Two independent accumulation chains: the stage's two 32-row MFMA
                // tiles, or (32-row stages) the even and odd K steps of one tile.
                constexpr bool kPair = RT >= 64;
                f32x16 acc0, acc1;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float4 n0 = *reinterpret_cast<const float4*>(curn + sp * 32 + i * 8 + h * 4);
                    acc0[i * 4 + 0] = n0.x; acc0[i * 4 + 1] = n0.y; acc0[i * 4 + 2] = n0.z; acc0[i * 4 + 3] = n0.w;
                    if (kPair) {
                        const float4 n1 = *reinterpret_cast<const float4*>(curn + (sp + 1) * 32 + i * 8 + h * 4);
                        acc1[i * 4 + 0] = n1.x; acc1[i * 4 + 1] = n1.y; acc1[i * 4 + 2] = n1.z; acc1[i * 4 + 3] = n1.w;
                    } else {
                        acc1[i * 4 + 0] = 0.0f; acc1[i * 4 + 1] = 0.0f; acc1[i * 4 + 2] = 0.0f; acc1[i * 4 + 3] = 0.0f;
                    }
                }
                const unsigned char* rowA0 = cur + (sp * 32 + col) * row_bytes;
                const unsigned char* rowA1 = kPair ? cur + ((sp + 1) * 32 + col) * row_bytes : rowA0;
                if (BF16) {
                    const bf16x8* a0 = reinterpret_cast<const bf16x8*>(rowA0 + h * KP);
                    const bf16x8* a1 = reinterpret_cast<const bf16x8*>(rowA1 + h * KP);
#pragma unroll
                    for (int s = 0; s < MAXSTEPS; s++)
                        if (s < steps) {
                            if (kPair) {
                                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0[s], bq[s], acc0, 0, 0, 0);
                                acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[s], bq[s], acc1, 0, 0, 0);
                            } else if (s & 1) {
                                acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0[s], bq[s], acc1, 0, 0, 0);
                            } else {
                                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0[s], bq[s], acc0, 0, 0, 0);
                            }
                        }
                } else {
                    // K steps come in fours (KP is a multiple of 8): one 16-byte LDS
                    // read feeds four MFMAs
                    const float4* a0 = reinterpret_cast<const float4*>(rowA0) + h * (KP / 8);
                    const float4* a1 = reinterpret_cast<const float4*>(rowA1) + h * (KP / 8);
#pragma unroll
                    for (int s4 = 0; s4 < MAXSTEPS / 4; s4++)
                        if (s4 * 4 < steps) {
                            const float4 x0 = a0[s4];
                            const float4 x1 = kPair ? a1[s4] : x0;
                            const float v0[4] = {x0.x, x0.y, x0.z, x0.w};
                            const float v1[4] = {x1.x, x1.y, x1.z, x1.w};
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                const int s = s4 * 4 + u;
                                if (kPair) {
                                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(v0[u], bqf[s], acc0, 0, 0, 0);
                                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(v1[u], bqf[s], acc1, 0, 0, 0);
                                } else if (u & 1) {
                                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(v0[u], bqf[s], acc1, 0, 0, 0);
                                } else {
                                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(v0[u], bqf[s], acc0, 0, 0, 0);
                                }
                            }
                        }
                }
                if (!kPair) {
#pragma unroll
                    for (int r = 0; r < 16; r++) acc0[r] += acc1[r];
                }
                // the query's own row: only the tile at the wave's query rows holds
                // it (a wave-uniform test), at offset col of the lane's query
                float gm[4], m;
                group_max(acc0, gm, m);
                epilogue<E, QC, PR>(acc0, gm, m, r0 + sp * 32 == qbase, r0 + sp * 32, h, col, lane,
                                    thr, qn, lists, myq, a.rowmap, floor);
                if (kPair) {
                    group_max(acc1, gm, m);
                    epilogue<E, QC, PR>(acc1, gm, m, r0 + (sp + 1) * 32 == qbase, r0 + (sp + 1) * 32,
                                        h, col, lane, thr, qn, lists, myq, a.rowmap, floor);
                }
            }
        }
        // stage st+1's copies have landed in LDS before any wave passes the barrier
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    };
    for (uint32_t st = 0; st < nst; st++) {
        if (PR && st > 0 && (st & 63u) == 0u) {
            idc = idn;
            idn = load_ids(st + 64);
        }
        if (st & 1u) step(tile1, negh1, tile0, negh0, st);
        else step(tile0, negh0, tile1, negh1, st);
    }
    // final flush of every non-empty queue
    const uint64_t left = __ballot(qn > 0);
    uint64_t need = (left | (left >> 32)) & 0xFFFFFFFFull;
    while (need) {
        int bs[kMergeBatch];
        const int nb = pop_queries(need, bs);
        uint64_t tk[kMergeBatch];
        merge_queries<E, QC>(bs, nb, qn, lane, lists, myq, PR ? a.rowmap : nullptr, tk);
    }
}

// join the per-split sorted lists of each query (one wave per query)
template <int E>
__global__ __launch_bounds__(256) void k_merge_splits(uint64_t* __restrict__ lists, uint32_t npad,
                                                      int splits) {
    constexpr int S = 64 * E;
    const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= npad) return;
    uint64_t v[E];
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = lists[(size_t)q * S + e * 64 + lane];
    for (int sp = 1; sp < splits; sp++) {
        const uint64_t* other = lists + ((size_t)sp * npad + q) * S;
        // merge the other list in chunks of 64 (each chunk sorted, ascending)
#pragma unroll
        for (int e = 0; e < E; e++) {
            uint64_t qv = other[e * 64 + lane];
            wave_merge<E>(v, qv);
        }
    }
#pragma unroll
    for (int e = 0; e < E; e++) lists[(size_t)q * S + e * 64 + lane] = v[e];
}

// Exact FP32 re-rank of the shortlist; one wave per query, lane = candidate.
// The wave first stages its 64 candidate rows into LDS with coalesced row
// loads (one row per load instruction); each lane then accumulates its own
// distance in dimension order with fmaf, the oracle's arithmetic.  Rows are
// padded to d+1 floats so the lanes' column reads fall in distinct banks.
constexpr int kRerankWaves = 4;
template <int E>
__global__ __launch_bounds__(kRerankWaves * 64) void k_rerank(const float* __restrict__ x,
                                                              uint32_t n, int d,
                                                              const uint64_t* __restrict__ lists,
                                                              int k, int32_t* __restrict__ out_idx,
                                                              float* __restrict__ out_dist) {
    extern __shared__ float rsh[];  // per wave: 64 rows of d+1 floats, then the query row
    constexpr int S = 64 * E;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t q = blockIdx.x * kRerankWaves + wave;
    if (q >= n) return;  // whole waves leave; nothing below is block-wide
    const int ld = d + 1;
    float* rows = rsh + (size_t)wave * (64 * ld + d);
    float* xq = rows + 64 * ld;
    for (int t = lane; t < d; t += 64) xq[t] = x[(size_t)q * d + t];
    const uint64_t* ql = lists + (size_t)q * S;  // the query's list
    uint64_t key[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint64_t lk = ql[e * 64 + lane];
        const uint32_t c = (uint32_t)lk;
        const bool ok = lk != kInfKey;
        for (int r = 0; r < 64; r++) {
            const uint32_t cr = __builtin_amdgcn_readlane(ok ? c : 0xFFFFFFFFu, r);
            if (cr == 0xFFFFFFFFu) continue;  // wave-uniform: an empty slot
            for (int t = lane; t < d; t += 64) rows[r * ld + t] = x[(size_t)cr * d + t];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (!ok) {
            key[e] = kInfKey;
        } else {
            const float* xc = rows + lane * ld;
            float s = 0.0f;
            for (int t = 0; t < d; t++) {
                const float df = xq[t] - xc[t];
                s = fmaf(df, df, s);
            }
            key[e] = ((uint64_t)__float_as_uint(s) << 32) | c;
        }
        __builtin_amdgcn_wave_barrier();  // rows are reused by the next e
    }
    wave_sort<E>(key);
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int g = e * 64 + lane;
        if (g < k) {
            const bool ok = key[e] != kInfKey;
            out_idx[(size_t)q * k + g] = ok ? (int32_t)(uint32_t)key[e] : -1;
            out_dist[(size_t)q * k + g] = ok ? __uint_as_float((uint32_t)(key[e] >> 32)) : INFINITY;
        }
    }
}

// The re-rank of the pruned sweep: lists and xs (X in cell order, rows of dp
// floats, dp a multiple of 4) are in position order.  Wave w takes position w
// (the caller's row rowmap[w]; none for padding); lane = candidate, reading its
// own row xs[posmap[c]] with 16-byte loads.  A query's candidates lie mostly
// in its own cell, so those rows are L2 hits, and without LDS staging many
// waves are resident to hide the remaining latency.  The distance is the
// oracle's fmaf over t in order, as in k_rerank.
template <int E>
__global__ __launch_bounds__(256) void k_rerank_cells(const float* __restrict__ xs, int dp, int d,
                                                      uint32_t npos,
                                                      const uint64_t* __restrict__ lists,
                                                      const uint32_t* __restrict__ rowmap,
                                                      const uint32_t* __restrict__ posmap, int k,
                                                      int32_t* __restrict__ out_idx,
                                                      float* __restrict__ out_dist) {
    __shared__ float xqs[4][128];  // d <= 128
    constexpr int S = 64 * E;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t w = blockIdx.x * 4 + wave;
    if (w >= npos) return;  // whole waves leave; nothing below is block-wide
    const uint32_t q = rowmap[w];
    if (q == 0xFFFFFFFFu) return;
    float* xq = xqs[wave];
    for (int t = lane; t < d; t += 64) xq[t] = xs[(size_t)w * dp + t];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int d4 = d & ~3;
    uint64_t key[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint64_t lk = lists[(size_t)w * S + e * 64 + lane];
        key[e] = kInfKey;
        if (lk != kInfKey) {
            const uint32_t c = (uint32_t)lk;
            const float* row = xs + (size_t)posmap[c] * dp;
            const float4* r4 = reinterpret_cast<const float4*>(row);
            float acc = 0.0f;
            for (int t = 0; t < d4; t += 4) {
                const float4 v = r4[t >> 2];
                float df = xq[t] - v.x;
                acc = fmaf(df, df, acc);
                df = xq[t + 1] - v.y;
                acc = fmaf(df, df, acc);
                df = xq[t + 2] - v.z;
                acc = fmaf(df, df, acc);
                df = xq[t + 3] - v.w;
                acc = fmaf(df, df, acc);
            }
            for (int t = d4; t < d; t++) {
                const float df = xq[t] - row[t];
                acc = fmaf(df, df, acc);
            }
            key[e] = ((uint64_t)__float_as_uint(acc) << 32) | c;
        }
    }
    wave_sort<E>(key);
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int g = e * 64 + lane;
        if (g < k) {
            const bool ok = key[e] != kInfKey;
            out_idx[(size_t)q * k + g] = ok ? (int32_t)(uint32_t)key[e] : -1;
            out_dist[(size_t)q * k + g] = ok ? __uint_as_float((uint32_t)(key[e] >> 32)) : INFINITY;
        }
    }
}

// ------------------------------------------------- pruned sweep (large n)
//
// For large n the database is first put in cell order: rows are assigned to
// the nearest of m seed rows, stably sorted by cell, and each cell is padded
// to whole 128-row query blocks (padding rows score -inf).  Every 64- or
// 128-row group then has a centroid c and radius r (in the packed, rounded
// coordinates the MFMA scores use).  Pass 1 sweeps, for each query block, the
// stages of its own cell.  If a query's list is then full, its S-th distance
// D_q bounds every candidate that can still enter it; a stage whose lower
// bound |c_b - c_s| - r_b - r_s exceeds sqrt(max_q D_q) (with a margin that
// covers the FP32 accumulation error of the scores) can hold no candidate for
// any query of the block and is skipped.  Pass 2 sweeps the remaining stages.
// Skipped rows would never have entered a list, and the threshold compares
// (score, row) keys, so the lists equal those of the full sweep.

constexpr uint32_t kBlockRows = 128;  // queries per workgroup (4 waves x 32)
constexpr uint32_t kCellRows = 1024;  // rows per cell on average
constexpr uint32_t kMaxCells = 1024;
constexpr int kListBlocks = 8;        // query blocks per workgroup of k_stage_lists
constexpr uint32_t kPruneMinRows = 65536;  // smaller problems sweep every stage

__device__ inline float packed_at(const void* P, int KP, uint32_t row, int t, bool bf16) {
    if (t >= KP) return 0.0f;
    return bf16 ? bf16_to_f32(reinterpret_cast<const u16*>(P)[(size_t)row * KP + t])
                : reinterpret_cast<const float*>(P)[(size_t)row * KP + t];
}

// Seed c = packed row floor(c n / m), as BF16 (K padded to KS, a multiple of
// 16), and its -0.5 |s|^2 from the rounded values; padding seeds up to mpad
// (a multiple of 32) score -inf.
template <bool BF16>
__global__ __launch_bounds__(64) void k_seed_rows(const void* __restrict__ P, int KP, int KS,
                                                  uint32_t n, uint32_t m,
                                                  u16* __restrict__ seeds,
                                                  float* __restrict__ snegh) {
    const uint32_t c = blockIdx.x;  // < mpad
    const int lane = threadIdx.x;
    const uint32_t row = c < m ? (uint32_t)(((uint64_t)c * n) / m) : 0u;
    float part = 0.0f;
    for (int t = lane; t < KS; t += 64) {
        const float v = c < m ? packed_at(P, KP, row, t, BF16) : 0.0f;
        const u16 h = f32_to_bf16_rne(v);
        seeds[(size_t)c * KS + t] = h;
        const float r = bf16_to_f32(h);
        part += r * r;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) part += __shfl_xor(part, o, 64);
    if (lane == 0) snegh[c] = c < m ? -0.5f * part : -INFINITY;
}

// cell[i] = the seed of largest score p_i . s - 0.5 |s|^2 (the nearest; the
// smallest seed on ties), count[c] = rows per cell.  A BF16 MFMA GEMM with an
// argmax epilogue: one wave per 32 rows (the B operand, as in the shortlist),
// 32 seeds per A tile read straight from the L2-resident seed matrix.  The
// cells only group rows for the pruned sweep, whose result does not depend on
// them, so BF16 rounding here is harmless.
template <bool BF16>
__global__ __launch_bounds__(256) void k_assign_cells(const void* __restrict__ P, int KP, int KS,
                                                      uint32_t n, const u16* __restrict__ seeds,
                                                      const float* __restrict__ snegh,
                                                      uint32_t mpad, uint32_t* __restrict__ cell,
                                                      uint32_t* __restrict__ count,
                                                      const float* __restrict__ negh,
                                                      float* __restrict__ dist2,
                                                      double* __restrict__ cell_d2) {
    const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
    const uint32_t i = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 32 + col;  // this lane's row
    const int steps = KS / 16;
    bf16x8 bq[8];
#pragma unroll
    for (int st = 0; st < 8; st++) {
        if (st >= steps) break;
        union {
            u16 h16[8];
            bf16x8 v;
        } f;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int t = h * (KS / 2) + st * 8 + e;
            f.h16[e] = f32_to_bf16_rne(i < n ? packed_at(P, KP, i, t, BF16) : 0.0f);
        }
        bq[st] = f.v;
    }
    float best = -INFINITY;
    uint32_t bi = 0;
    for (uint32_t c0 = 0; c0 < mpad; c0 += 32) {
        f32x16 acc;
#pragma unroll
        for (int g = 0; g < 4; g++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[g * 4 + j] = snegh[c0 + g * 8 + h * 4 + j];
        const bf16x8* arow =
            reinterpret_cast<const bf16x8*>(seeds + (size_t)(c0 + col) * KS + h * (KS / 2));
#pragma unroll
        for (int st = 0; st < 8; st++) {
            if (st >= steps) break;
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(arow[st], bq[st], acc, 0, 0, 0);
        }
#pragma unroll
        for (int g = 0; g < 4; g++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (acc[g * 4 + j] > best) {
                    best = acc[g * 4 + j];
                    bi = c0 + g * 8 + h * 4 + j;
                }
    }
    // the two lanes of a row hold the two halves of each seed tile's rows
    const float ob = __shfl_xor(best, 32, 64);
    const uint32_t oi = __shfl_xor(bi, 32, 64);
    if (ob > best || (ob == best && oi < bi)) bi = oi;
    if (h == 0 && i < n) {
        cell[i] = bi;
        atomicAdd(&count[bi], 1u);
        // squared distance to the seed, |p|^2 - 2 score (BF16 scores: a rank, not a distance)
        const float d2 = fmaxf(-2.0f * negh[i] - 2.0f * fmaxf(best, ob), 0.0f);
        dist2[i] = d2;
        atomicAdd(&cell_d2[bi], (double)d2);
    }
}

// Sort key of row i: its cell, then 1 for a far row (squared distance to the
// seed above twice the cell's mean), so each cell's far rows (outliers,
// uniform background) go to its last blocks and its other blocks stay compact.
__global__ __launch_bounds__(256) void k_cell_keys(const uint32_t* __restrict__ cell,
                                                   const float* __restrict__ dist2,
                                                   const double* __restrict__ cell_d2,
                                                   const uint32_t* __restrict__ count, uint32_t n,
                                                   uint32_t* __restrict__ key) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = cell[i];
    const double mean = cell_d2[c] / (double)count[c];
    key[i] = (c << 1) | ((double)dist2[i] > 2.0 * mean ? 1u : 0u);
}

// position of sorted row r: its cell's padded start plus its rank in the cell
__global__ __launch_bounds__(256) void k_cell_layout(const uint32_t* __restrict__ cell_sorted,
                                                     const uint32_t* __restrict__ order, uint32_t n,
                                                     const uint32_t* __restrict__ start,
                                                     const uint32_t* __restrict__ pstart,
                                                     uint32_t* __restrict__ rowmap,
                                                     uint32_t* __restrict__ posmap) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const uint32_t c = cell_sorted[r] >> 1;  // sort key: cell, then far rows last
    const uint32_t p = pstart[c] + (r - start[c]);
    const uint32_t i = order[r];
    rowmap[p] = i;
    posmap[i] = p;
}

// packed rows in cell order (padding rows zero, score -inf)
__global__ __launch_bounds__(256) void k_gather_rows(const unsigned char* __restrict__ P,
                                                     int row_bytes, const float* __restrict__ negh,
                                                     const uint32_t* __restrict__ rowmap,
                                                     uint32_t npos, unsigned char* __restrict__ P2,
                                                     float* __restrict__ negh2) {
    const int cpr = row_bytes / 16;
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (uint64_t)npos * cpr) return;
    const uint32_t p = (uint32_t)(t / cpr);
    const int o = (int)(t - (uint64_t)p * cpr);
    const uint32_t src = rowmap[p];
    u32x4 v = {0u, 0u, 0u, 0u};
    if (src != 0xFFFFFFFFu) v = reinterpret_cast<const u32x4*>(P + (size_t)src * row_bytes)[o];
    reinterpret_cast<u32x4*>(P2 + (size_t)p * row_bytes)[o] = v;
    if (o == 0) negh2[p] = src != 0xFFFFFFFFu ? negh[src] : -INFINITY;
}

// X rows in cell order, dp >= d floats per row (padding rows and columns zero)
__global__ __launch_bounds__(256) void k_gather_x(const float* __restrict__ x, int d, int dp,
                                                  const uint32_t* __restrict__ rowmap, uint32_t npos,
                                                  float* __restrict__ xs) {
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (uint64_t)npos * dp) return;
    const uint32_t p = (uint32_t)(t / dp);
    const int f = (int)(t - (uint64_t)p * dp);
    const uint32_t src = rowmap[p];
    xs[t] = (src != 0xFFFFFFFFu && f < d) ? x[(size_t)src * d + f] : 0.0f;
}

// Bounds of each G-row group in cell order (one wavefront per group; lane t
// holds dimensions t and t + 64): centroid, radius (largest distance of a row
// to it), largest squared row norm, row count.
template <bool BF16>
__global__ __launch_bounds__(256) void k_group_bounds(const void* __restrict__ P2, int KP,
                                                      const uint32_t* __restrict__ rowmap, int G,
                                                      uint32_t ngroups, float* __restrict__ cen,
                                                      float* __restrict__ rad,
                                                      float* __restrict__ gnorm,
                                                      uint32_t* __restrict__ cnt) {
    const uint32_t g = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (g >= ngroups) return;  // whole wavefronts
    float c0 = 0.0f, c1 = 0.0f;
    uint32_t k = 0;
    for (int r = 0; r < G; r++) {
        const uint32_t p = g * (uint32_t)G + r;
        if (rowmap[p] == 0xFFFFFFFFu) continue;  // uniform
        k++;
        c0 += packed_at(P2, KP, p, lane, BF16);
        c1 += packed_at(P2, KP, p, lane + 64, BF16);
    }
    float r2 = 0.0f, n2 = 0.0f;
    if (k) {
        c0 /= (float)k;
        c1 /= (float)k;
        for (int r = 0; r < G; r++) {
            const uint32_t p = g * (uint32_t)G + r;
            if (rowmap[p] == 0xFFFFFFFFu) continue;
            const float v0 = packed_at(P2, KP, p, lane, BF16), v1 = packed_at(P2, KP, p, lane + 64, BF16);
            float dd = (v0 - c0) * (v0 - c0) + (v1 - c1) * (v1 - c1);
            float nn = v0 * v0 + v1 * v1;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                dd += __shfl_xor(dd, o, 64);
                nn += __shfl_xor(nn, o, 64);
            }
            r2 = fmaxf(r2, dd);
            n2 = fmaxf(n2, nn);
        }
    }
    if (lane < KP) cen[(size_t)g * KP + lane] = c0;
    if (lane + 64 < KP) cen[(size_t)g * KP + lane + 64] = c1;
    if (lane == 0) {
        rad[g] = sqrtf(r2);
        gnorm[g] = n2;
        cnt[g] = k;
    }
}

// Pass-1 range of each query block: the stages of its own cell.
__global__ __launch_bounds__(256) void k_block_ranges(const uint32_t* __restrict__ pstart,
                                                      uint32_t m, uint32_t nblk, uint32_t RT,
                                                      uint2* __restrict__ range) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nblk) return;
    const uint32_t p = b * kBlockRows;
    if (p >= pstart[m]) {  // padding past the last cell
        range[b] = make_uint2(0u, 0u);
        return;
    }
    uint32_t lo = 0, hi = m;  // pstart[lo] <= p < pstart[hi]; empty cells are passed over
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (pstart[mid] <= p) lo = mid;
        else hi = mid;
    }
    range[b] = make_uint2(pstart[lo] / RT, (pstart[lo + 1] - pstart[lo]) / RT);
}

// Pass-1 floors.  For each query, a score that at least S rows of its own
// cell (itself excluded) reach: any such value is at most the S-th best score
// of the cell, and so of the whole database, so every row of the query's
// final list still passes it, and pass 1, starting from it, queues only the
// rows at or above it (about a third of what it queued from -inf, where the
// threshold climbs from nothing).  The scores are the shortlist kernel's bits:
// the same bias, the same MFMA chain over the K steps in order, the query's
// own row at -inf.  One sweep over the cell counts each query's rows in
// kFloorBins bins of squared distance d2 = |x|^2 - 2 score on a log scale
// (sixteen bins per octave, from r2 / 8 to 2 r2, r2 the block's squared
// radius, rows beyond them in the end bins), in LDS; the floor is the score
// at the far edge of the bin after the one where the count from the nearest
// reaches S (a bin, at least 3% of d2, above any rounding), or -inf if no
// such bin lies within the window or the floor would lose its meaning to
// cancellation against |x|^2.  One workgroup per pass-1 work item (a
// 128-query block; items come largest cell first, so no long block starts
// last), wave w taking queries w*32 .. w*32 + 31 as in the shortlist kernel;
// rows are read from global memory (the cell is a few hundred KB and stays
// in L2), one tile ahead of their MFMA chain.
constexpr int kFloorBins = 64;
constexpr int kFloorPerOct = 16;  // bins per octave of d2
constexpr int kFloorBelow = 48;   // bins below r2 (three octaves); one octave above
template <int STEPS>
__global__ __launch_bounds__(256) void k_cell_floor(const u16* __restrict__ P,
                                                    const float* __restrict__ negh,
                                                    const uint2* __restrict__ range,
                                                    const float* __restrict__ brad,
                                                    const uint4* __restrict__ items, int KP,
                                                    int RT, int S, float* __restrict__ floor_out) {
    __shared__ uint32_t hist[4][kFloorBins][kQPerWave];  // bank = query
    const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t blk = __builtin_amdgcn_readfirstlane(items[blockIdx.x].x);
    const uint2 rg = range[blk];
    const uint32_t qbase = (blk * 4 + wave) * kQPerWave;
    const uint32_t q = qbase + col;
    bf16x8 bq[STEPS];
    {
        const bf16x8* src = reinterpret_cast<const bf16x8*>(P + (size_t)q * KP + h * (KP / 2));
#pragma unroll
        for (int s = 0; s < STEPS; s++) bq[s] = src[s];
    }
    const uint32_t r0 = __builtin_amdgcn_readfirstlane(rg.x * (uint32_t)RT);
    const uint32_t r1 = __builtin_amdgcn_readfirstlane((rg.x + rg.y) * (uint32_t)RT);
    const float xn2 = -2.0f * negh[q];  // |x|^2 of the query's rounded coordinates
    const float r2 = brad[blk] * brad[blk];
    const bool scaled = r2 > 0.0f && r2 < INFINITY && xn2 < INFINITY;
    const float inv_r2 = scaled ? 1.0f / r2 : 0.0f;
    uint32_t* hq = &hist[wave][0][0];
    for (int i = lane; i < kQPerWave * kFloorBins; i += 64) hq[i] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    struct Ops {
        float4 nv[4];
        bf16x8 a[STEPS];
    };
    auto load = [&](uint32_t row0, Ops& o) {
#pragma unroll
        for (int i = 0; i < 4; i++) o.nv[i] = *reinterpret_cast<const float4*>(negh + row0 + i * 8 + h * 4);
        const bf16x8* ap =
            reinterpret_cast<const bf16x8*>(P + (size_t)(row0 + col) * KP + h * (KP / 2));
#pragma unroll
        for (int s = 0; s < STEPS; s++) o.a[s] = ap[s];
    };
    // bin of a score from the bits of d2 / r2 (its exponent and top four
    // mantissa bits: sixteen bins per octave), clamped to the window
    auto bin_of = [&](float sc) {
        const float d2 = fmaxf(xn2 - 2.0f * sc, 0.0f) * inv_r2;
        const int bits = (int)(__float_as_uint(d2) >> 19);  // exponent * 16 + top four mantissa bits
        return min(kFloorBins - 1, max(0, bits - (127 * kFloorPerOct - kFloorBelow)));
    };
    auto count = [&](uint32_t row0, const Ops& o) {
        f32x16 c;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            c[i * 4 + 0] = o.nv[i].x;
            c[i * 4 + 1] = o.nv[i].y;
            c[i * 4 + 2] = o.nv[i].z;
            c[i * 4 + 3] = o.nv[i].w;
        }
#pragma unroll
        for (int s = 0; s < STEPS; s++) c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(o.a[s], bq[s], c, 0, 0, 0);
        if (row0 == qbase) mask_self(c, h, col);
#pragma unroll
        for (int r = 0; r < 16; r++)
            if (scaled && c[r] != -INFINITY) atomicAdd(&hist[wave][bin_of(c[r])][col], 1u);
    };
    if (r0 < r1 && __ballot(scaled)) {
        // two operand sets take turns: one loads while the other's MFMAs run
        Ops o0, o1;
        load(r0, o0);
        for (uint32_t row0 = r0; row0 < r1; row0 += 64) {
            if (row0 + 32 < r1) load(row0 + 32, o1);
            count(row0, o0);
            if (row0 + 32 < r1) {
                if (row0 + 64 < r1) load(row0 + 64, o0);
                count(row0 + 32, o1);
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (h == 0) {
        float fl = -INFINITY;
        if (scaled) {
            uint32_t cum = 0;
            int b = 0;
            for (; b < kFloorBins; b++) {
                cum += hist[wave][b][col];
                if (cum >= (uint32_t)S) break;
            }
            // every row counted so far lies below the far edge of bin b, so
            // (a bin further) above the score at the far edge of bin b + 1
            if (b + 2 < kFloorBins) {
                const int e = b + 2 - kFloorBelow;
                const float edge = ldexpf(r2, e >> 4) * (1.0f + (float)(e & 15) / 16.0f);
                if (edge >= 1e-4f * xn2) fl = 0.5f * (xn2 - edge);
            }
        }
        floor_out[q] = fl;
    }
}

// Pass-2 stage lists.  A workgroup takes kListBlocks query blocks; each
// thread tests one stage against all of them per round.  A block's bound D is
// the largest distance^2 threshold of its queries after pass 1 (|x|^2 - 2 thr,
// +inf while a list is not full).  WRITE = false counts the kept stages per
// block; WRITE = true writes them, in ascending order, at offs[b] (never more
// than the count pass found).
template <bool WRITE>
__global__ __launch_bounds__(256) void k_stage_lists(
    const float* __restrict__ bcen, const float* __restrict__ brad,
    const float* __restrict__ bnorm, const uint32_t* __restrict__ bcnt,
    const float* __restrict__ scen, const float* __restrict__ srad,
    const float* __restrict__ snorm, const uint32_t* __restrict__ scnt,
    const uint2* __restrict__ own, const float* __restrict__ negh2,
    const uint64_t* __restrict__ lists, int S, int KP, uint32_t nblk, uint32_t nst, int all,
    uint32_t* __restrict__ counts, const uint32_t* __restrict__ offs, uint32_t* __restrict__ ids) {
    // WRITE reads counts (the count pass's totals) as a bound; all: every
    // stage outside the block's own cell, no bound (the unpruned sweep)
    __shared__ __align__(16) float cb[kListBlocks][128];
    __shared__ float dmax[kListBlocks];
    __shared__ uint32_t wcnt[kListBlocks][4];
    __shared__ uint32_t total[kListBlocks];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t b0 = blockIdx.x * kListBlocks;
    if (tid < kListBlocks) {
        dmax[tid] = 0.0f;
        total[tid] = 0u;
    }
    for (int e = tid; e < kListBlocks * 128; e += 256) {
        const uint32_t b = b0 + e / 128;
        const int t = e % 128;
        cb[e / 128][t] = (b < nblk && t < KP) ? bcen[(size_t)b * KP + t] : 0.0f;
    }
    __syncthreads();
    // D of each block: 512 queries, two per thread; max of non-negative floats
    // by their bit patterns
    for (int e = tid; e < kListBlocks * (int)kBlockRows; e += 256) {
        const uint32_t b = b0 + e / kBlockRows;
        if (b >= nblk) continue;
        const uint32_t q = b * kBlockRows + e % kBlockRows;
        const float nh = negh2[q];
        if (nh == -INFINITY) continue;  // padding row
        const float thr = key_score(lists[(size_t)q * S + S - 1]);
        float dq = thr == -INFINITY ? INFINITY : fmaxf(-2.0f * nh - 2.0f * thr, 0.0f);
        atomicMax(reinterpret_cast<uint32_t*>(&dmax[e / kBlockRows]), __float_as_uint(dq));
    }
    __syncthreads();
    float D[kListBlocks], rb[kListBlocks], nb[kListBlocks];
    uint2 ob[kListBlocks];
    bool vb[kListBlocks];
#pragma unroll
    for (int j = 0; j < kListBlocks; j++) {
        const uint32_t b = b0 + j;
        vb[j] = b < nblk && bcnt[b] > 0;
        D[j] = all ? INFINITY : dmax[j];
        rb[j] = vb[j] ? brad[b] : 0.0f;
        nb[j] = vb[j] ? sqrtf(bnorm[b]) : 0.0f;
        ob[j] = vb[j] ? own[b] : make_uint2(0u, 0u);
    }
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t s0 = 0; s0 < nst; s0 += 256) {
        const uint32_t sidx = s0 + tid;
        const bool live = sidx < nst && scnt[sidx] > 0;
        float d2[kListBlocks];
#pragma unroll
        for (int j = 0; j < kListBlocks; j++) d2[j] = 0.0f;
        float rs = 0.0f, ns = 0.0f;
        if (live) {
            const float4* cs = reinterpret_cast<const float4*>(scen + (size_t)sidx * KP);
            for (int t4 = 0; t4 < KP / 4; t4++) {
                const float4 v = cs[t4];
#pragma unroll
                for (int j = 0; j < kListBlocks; j++) {
                    const float4 w = *reinterpret_cast<const float4*>(&cb[j][4 * t4]);
                    const float e0 = w.x - v.x, e1 = w.y - v.y, e2 = w.z - v.z, e3 = w.w - v.w;
                    d2[j] += e0 * e0 + e1 * e1 + e2 * e2 + e3 * e3;
                }
            }
            rs = srad[sidx];
            ns = sqrtf(snorm[sidx]);
        }
#pragma unroll
        for (int j = 0; j < kListBlocks; j++) {
            bool keep = live && vb[j] && !(sidx >= ob[j].x && sidx < ob[j].x + ob[j].y);
            if (keep) {
                const float scale = nb[j] + ns;
                const float lb = sqrtf(d2[j]) * (1.0f - 1e-5f) - (rb[j] + rs) * (1.0f + 1e-5f) -
                                 1e-6f * scale;
                const float margin = 2e-4f * scale * scale + 1e-5f * D[j];
                keep = !(lb > 0.0f && lb * lb > D[j] + margin);
            }
            const uint64_t bal = __ballot(keep);
            if (!WRITE) {
                if (lane == 0 && bal) atomicAdd(&total[j], (uint32_t)__popcll(bal));
            } else {
                if (lane == 0) wcnt[j][wave] = (uint32_t)__popcll(bal);
                __syncthreads();
                uint32_t before = total[j];
                for (int w = 0; w < wave; w++) before += wcnt[j][w];
                const uint32_t at = before + (uint32_t)__popcll(bal & below);
                if (keep && at < counts[b0 + j]) ids[offs[b0 + j] + at] = sidx;
                __syncthreads();
                if (tid == 0) total[j] += wcnt[j][0] + wcnt[j][1] + wcnt[j][2] + wcnt[j][3];
                __syncthreads();
            }
        }
    }
    if (!WRITE) {
        __syncthreads();
        if (tid < kListBlocks && b0 + tid < nblk) counts[b0 + tid] = total[tid];
    }
}

// floor[q] = the score of query position q's S-th list entry (a pass's start)
__global__ __launch_bounds__(256) void k_list_floor(const uint64_t* __restrict__ lists, int S,
                                                    uint32_t npos, float* __restrict__ floor) {
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q < npos) floor[q] = key_score(lists[(size_t)q * S + S - 1]);
}

// Fold the aux lists of split work items into their block's lists: merges[m]
// = (block, first aux slot, aux count); one wavefront per query of the block,
// merging the slots in order (disjoint candidates, so the result is the S
// smallest keys of the union).
template <int E>
__global__ __launch_bounds__(256) void k_merge_aux(uint64_t* __restrict__ lists,
                                                   const uint64_t* __restrict__ aux_lists,
                                                   const uint4* __restrict__ merges,
                                                   uint32_t n_merges) {
    constexpr int S = 64 * E;
    const uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= n_merges * kBlockRows) return;  // whole wavefronts
    const uint4 mg = merges[w / kBlockRows];
    const uint32_t j = w % kBlockRows;
    uint64_t* dst = lists + ((size_t)mg.x * kBlockRows + j) * S;
    uint64_t v[E];
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = dst[e * 64 + lane];
    for (uint32_t a = 0; a < mg.z; a++) {
        const uint64_t* src = aux_lists + ((size_t)(mg.y + a) * kBlockRows + j) * S;
#pragma unroll
        for (int e = 0; e < E; e++) wave_merge<E>(v, src[e * 64 + lane]);
    }
#pragma unroll
    for (int e = 0; e < E; e++) dst[e * 64 + lane] = v[e];
}

__global__ __launch_bounds__(256) void k_range_counts(const uint2* __restrict__ range,
                                                      uint32_t nblk, uint32_t* __restrict__ counts) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b < nblk) counts[b] = range[b].y;
}

// the stages of block b's range, written at offs[b] (one workgroup per block)
__global__ __launch_bounds__(256) void k_range_ids(const uint2* __restrict__ range,
                                                   const uint32_t* __restrict__ offs,
                                                   uint32_t* __restrict__ ids) {
    const uint2 r = range[blockIdx.x];
    const uint32_t o = offs[blockIdx.x];
    for (uint32_t t = threadIdx.x; t < r.y; t += 256) ids[o + t] = r.x + t;
}


size_t rerank_lds_bytes(int d) {
    return (size_t)kRerankWaves * (64 * (d + 1) + d) * sizeof(float);
}

size_t shortlist_lds_bytes(bool bf16, int KP, int RT) {
    const int elem = bf16 ? 2 : 4;
    const size_t row_bytes = (size_t)KP * elem + 16;
    return 2 * RT * row_bytes + 2 * RT * sizeof(float);
}

constexpr size_t kLdsCu = 160 * 1024;  // LDS per CU on gfx950

struct Shape {
    int W, RT, QC;  // waves per workgroup, rows per stage, queue capacity
};

// Candidate shapes, deepest stage first: the first whose two stage buffers
// leave LDS for two workgroups per CU (else one) and whose stage fits the
// staging registers is used.  Deeper stages mean fewer barriers per MFMA.
constexpr Shape kShapes[] = {{4, 128, 64}, {4, 64, 64}, {4, 32, 64}};

Shape shortlist_shape(bool bf16, int KP) {
    for (int per_cu : {2, 1})
        for (const Shape& sh : kShapes) {
            const size_t lds = shortlist_lds_bytes(bf16, KP, sh.RT);
            const int chunks = sh.RT * (KP * (bf16 ? 2 : 4) / 16);
            if (lds * per_cu <= kLdsCu && chunks <= kStageChunks * kWavesMax * 64) return sh;
        }
    throw Error(MRS_INVALID_ARGUMENT, "d too large for the shortlist kernel");
}

}  // namespace

void knn_exhaustive_dev(mrs_ctx* ctx, const float* x, int64_t n64, int32_t d, int32_t k,
                        int32_t precision, int32_t shortlist, int32_t* out_idx, float* out_dist,
                        KnnTimings* timings, const KnnScratch* scratch) {
    if (n64 <= 1) throw Error(MRS_NO_DATA, "kNN needs at least two points (NoData)");
    if (k <= 0 || k > n64 - 1) throw Error(MRS_INVALID_ARGUMENT, "k must be in [1, n-1]");
    if (d <= 0 || d > 128) throw Error(MRS_INVALID_ARGUMENT, "exhaustive kNN supports 1 <= d <= 128");
    if (n64 > (int64_t)INT32_MAX - 256) throw Error(MRS_INVALID_ARGUMENT, "n too large");
    if (precision != MRS_KNN_FP32 && precision != MRS_KNN_BF16_RERANK)
        throw Error(MRS_INVALID_ARGUMENT, "unknown kNN precision");
    int want = std::max(shortlist, k);
    const int E = want <= 64 ? 1 : 2;
    if (want > 128) throw Error(MRS_INVALID_ARGUMENT, "shortlist must be <= 128");
    const bool bf16 = precision == MRS_KNN_BF16_RERANK;
    const uint32_t n = (uint32_t)n64;
    // bf16: K steps of 16 (8 per lane half); f32: steps of 2, rows kept 16-byte multiples
    const int KP = bf16 ? ((d + 15) / 16) * 16 : ((d + 7) / 8) * 8;
    // queries are padded to whole 8-wave blocks (a multiple of every shape's
    // block and of kSplitAlign)
    const uint32_t npad = ((n + 255) / 256) * 256;
    hipStream_t s = ctx->stream;

    if (scratch && timings) throw Error(MRS_INVALID_ARGUMENT, "no timings from scratch memory");
    // stage timing events, destroyed on every exit (including a throw); none
    // with caller scratch (nothing is created or waited for then)
    struct Events {
        hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Events() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } evs;
    hipEvent_t* ev = evs.e;
    if (!scratch)
        for (int i = 0; i < 4; i++) MRS_HIP(hipEventCreate(&ev[i]));
    auto mark = [&](int i) {
        if (ev[i]) MRS_HIP(hipEventRecord(ev[i], s));
    };
    mark(0);
    // caller scratch, carved in 256-byte aligned pieces
    size_t carved = 0;
    auto carve = [&](size_t bytes) -> void* {
        const size_t at = (carved + 255) / 256 * 256;
        if (at + bytes > scratch->bytes)
            throw Error(MRS_INVALID_ARGUMENT, "the kNN scratch is too small for this problem");
        carved = at + bytes;
        return scratch->base + at;
    };

    const size_t p_bytes = (size_t)npad * KP * (bf16 ? 2 : 4);
    DevBuf<unsigned char> P_own;
    DevBuf<float> negh_own;
    unsigned char* Pp;
    float* neghp;
    if (scratch) {
        Pp = static_cast<unsigned char*>(carve(p_bytes));
        neghp = static_cast<float*>(carve(sizeof(float) * npad));
    } else {
        P_own.alloc(p_bytes);
        negh_own.alloc(npad);
        Pp = P_own.p;
        neghp = negh_own.p;
    }
    if (bf16)
        hipLaunchKernelGGL(k_pack<true>, dim3(div_up(npad, 4)), dim3(256), 0, s, x, n, d, npad, KP,
                           (void*)Pp, neghp);
    else
        hipLaunchKernelGGL(k_pack<false>, dim3(div_up(npad, 4)), dim3(256), 0, s, x, n, d, npad,
                           KP, (void*)Pp, neghp);
    MRS_CHECK_LAUNCH();

    const Shape sh = shortlist_shape(bf16, KP);
    const int S = 64 * E;
    const size_t lds = shortlist_lds_bytes(bf16, KP, sh.RT);
    // one launch of the shortlist kernel compiled for this (precision, E,
    // shape); BF16 kernels are compiled per K-step count for the shapes that
    // take it (128-row stages: KP <= 64; 64-row stages: KP 80..128)
    auto run_shortlist = [&](const ShortlistArgs& sa, dim3 grid, bool pr) {
        bool launched = false;
        auto launch = [&](auto kern, int W) {
            ensure_dynamic_lds((const void*)kern, lds);
            hipLaunchKernelGGL(kern, grid, dim3(W * 64), lds, s, sa);
        };
        const int ksteps = bf16 ? KP / 16 : 0;
        auto pick = [&](auto b16, auto e, auto rt, auto qc, auto w) {
            constexpr bool B = decltype(b16)::value;
            constexpr int EE = decltype(e)::value, RTT = decltype(rt)::value;
            constexpr int QCC = decltype(qc)::value, WW = decltype(w)::value;
            if (launched || B != bf16 || EE != E || RTT != sh.RT || QCC != sh.QC || WW != sh.W) return;
            if constexpr (B) {
                auto st = [&](auto stc) {
                    constexpr int ST = decltype(stc)::value;
                    if (launched || ST != ksteps) return;
                    if (pr) launch(k_shortlist<B, EE, RTT, QCC, WW, ST, true>, WW);
                    else launch(k_shortlist<B, EE, RTT, QCC, WW, ST, false>, WW);
                    launched = true;
                };
                using std::integral_constant;
                if constexpr (RTT == 128) {
                    st(integral_constant<int, 1>{}); st(integral_constant<int, 2>{});
                    st(integral_constant<int, 3>{}); st(integral_constant<int, 4>{});
                } else if constexpr (RTT == 64) {
                    st(integral_constant<int, 5>{}); st(integral_constant<int, 6>{});
                    st(integral_constant<int, 7>{}); st(integral_constant<int, 8>{});
                }
            }
            if (!launched) {
                if (pr) launch(k_shortlist<B, EE, RTT, QCC, WW, 0, true>, WW);
                else launch(k_shortlist<B, EE, RTT, QCC, WW, 0, false>, WW);
                launched = true;
            }
        };
        using std::integral_constant;
        using T = integral_constant<bool, true>;
        using F = integral_constant<bool, false>;
        using I1 = integral_constant<int, 1>;
        using I2 = integral_constant<int, 2>;
        using I4 = integral_constant<int, 4>;
        using I32 = integral_constant<int, 32>;
        using I64 = integral_constant<int, 64>;
        using I128 = integral_constant<int, 128>;
        auto all = [&](auto rt, auto qc, auto w) {
            pick(T{}, I1{}, rt, qc, w);
            pick(T{}, I2{}, rt, qc, w);
            pick(F{}, I1{}, rt, qc, w);
            pick(F{}, I2{}, rt, qc, w);
        };
        all(I128{}, I64{}, I4{});
        all(I64{}, I64{}, I4{});
        all(I32{}, I64{}, I4{});
        if (!launched)
            throw Error(MRS_INVALID_ARGUMENT, "no shortlist kernel for this (precision, d) shape");
        MRS_CHECK_LAUNCH();
    };
    const size_t rlds = rerank_lds_bytes(d);
    auto rerank = [&](const uint64_t* lists) {
        auto go = [&](auto kern) {
            ensure_dynamic_lds((const void*)kern, rlds);
            hipLaunchKernelGGL(kern, dim3(div_up(n, kRerankWaves)), dim3(kRerankWaves * 64), rlds,
                               s, x, n, d, lists, k, out_idx, out_dist);
        };
        if (E == 1) go(k_rerank<1>);
        else go(k_rerank<2>);
        MRS_CHECK_LAUNCH();
    };
    int splits = 1;
    double swept = 1.0, list_mean = 0.0;
    uint32_t cells = 0, list_max = 0, aux_items = 0;

    // From kPruneMinRows rows the sweep runs in cell order: each query block
    // first sweeps its own cell, then (pruned) the stages its bound cannot rule
    // out, or (pruning off) every other stage.  The own cell fills the lists
    // with near-final thresholds, so the rest of the sweep, even when it visits
    // every stage, rarely inserts: most tiles leave on the max/ballot test.
    const bool by_cell = !scratch && n >= kPruneMinRows && sh.W * kQPerWave == (int)kBlockRows &&
                         ctx->knn_prune >= 0;
    const bool prune = by_cell && ctx->knn_prune > 0;
    if (!by_cell) {
        const uint32_t qblocks = npad / (uint32_t)(sh.W * kQPerWave);
        const uint32_t max_splits = scratch ? (uint32_t)std::max(1, scratch->max_splits) : 8u;
        // split count: forced, recorded for this shape class, tuned now, or the
        // built-in choice (enough query-block x split workgroups, 1024, to fill
        // the chip); the lists are the same for every split count
        auto layout = [&](uint32_t want, uint32_t* dbp) {
            const uint32_t sp = std::min(max_splits, std::max(1u, want));
            *dbp = ((npad + sp - 1) / sp + kSplitAlign - 1) / kSplitAlign * kSplitAlign;
            return (int)((npad + *dbp - 1) / *dbp);
        };
        const uint64_t shape = (uint64_t)bf16 | (uint64_t)E << 1 | (uint64_t)ceil_log2(npad) << 3 |
                               (uint64_t)KP << 9;
        int want = tune_forced(kTuneKnnSplits);
        if (want <= 0) want = tune_lookup(kTuneKnnSplits, shape);
        const bool tune = want <= 0 && !scratch && tune_enabled();
        if (want <= 0) want = (int)std::max<uint32_t>(1, (1024 + qblocks - 1) / qblocks);
        uint32_t db_per = 0, db_max = 0;
        splits = layout((uint32_t)want, &db_per);
        const int alloc_splits = tune ? layout(max_splits, &db_max) : splits;
        const size_t n_list = (size_t)alloc_splits * npad * S,
                     n_queue = (size_t)alloc_splits * npad * sh.QC;
        DevBuf<uint64_t> lists_own, queue_own;
        uint64_t *lists, *queue;
        if (scratch) {
            lists = static_cast<uint64_t*>(carve(sizeof(uint64_t) * n_list));
            queue = static_cast<uint64_t*>(carve(sizeof(uint64_t) * n_queue));
        } else {
            lists_own.alloc(n_list);
            queue_own.alloc(n_queue);
            lists = lists_own.p;
            queue = queue_own.p;
        }
        ShortlistArgs sa{};
        sa.P = Pp;
        sa.negh = neghp;
        sa.lists = lists;
        sa.queue = queue;
        sa.n = n;
        sa.npad = npad;
        sa.KP = KP;
        auto sweep = [&](int sp, uint32_t dbp) {
            sa.db_per_split = dbp;
            run_shortlist(sa, dim3(qblocks, sp), false);
            if (sp > 1) {
                if (E == 1)
                    hipLaunchKernelGGL(k_merge_splits<1>, dim3(div_up(npad, 4)), dim3(256), 0, s,
                                       lists, npad, sp);
                else
                    hipLaunchKernelGGL(k_merge_splits<2>, dim3(div_up(npad, 4)), dim3(256), 0, s,
                                       lists, npad, sp);
                MRS_CHECK_LAUNCH();
            }
        };
        if (tune) {
            // every distinct split count once, shortlist and merge timed with
            // events; the lists of the last run are the result (all runs give
            // the same lists)
            std::vector<int> cands;
            std::vector<uint32_t> dbps;
            for (uint32_t c : {1u, 2u, 4u, 8u}) {
                uint32_t dbp = 0;
                const int sp = layout(c, &dbp);
                if (std::find(cands.begin(), cands.end(), sp) == cands.end()) {
                    cands.push_back(sp);
                    dbps.push_back(dbp);
                }
            }
            std::vector<hipEvent_t> te(cands.size() + 1, nullptr);
            try {
                for (auto& e : te) MRS_HIP(hipEventCreate(&e));
                mark(1);
                MRS_HIP(hipEventRecord(te[0], s));
                for (size_t c = 0; c < cands.size(); c++) {
                    sweep(cands[c], dbps[c]);
                    MRS_HIP(hipEventRecord(te[c + 1], s));
                }
                mark(2);
                MRS_HIP(hipEventSynchronize(te.back()));
                std::vector<float> ms(cands.size());
                for (size_t c = 0; c < cands.size(); c++)
                    MRS_HIP(hipEventElapsedTime(&ms[c], te[c], te[c + 1]));
                splits = tune_record(kTuneKnnSplits, shape, cands, ms);
            } catch (...) {
                for (auto e : te)
                    if (e) (void)hipEventDestroy(e);
                throw;
            }
            for (auto e : te) (void)hipEventDestroy(e);
        } else {
            mark(1);
            sweep(splits, db_per);
            mark(2);
        }
        rerank(lists);
        mark(3);
        if (!scratch) MRS_HIP(hipEventSynchronize(ev[3]));  // before the buffers go
    } else {
        // ---- cell order (see the pruned-sweep kernels above)
        const int elem = bf16 ? 2 : 4;
        const uint32_t m = std::min<uint32_t>(kMaxCells, std::max<uint32_t>(1u, n / kCellRows));
        cells = m;
        const int KS = (KP + 15) / 16 * 16;
        const uint32_t mpad = (m + 31) / 32 * 32;
        DevBuf<u16> seeds((size_t)mpad * KS);
        DevBuf<float> snegh(mpad);
        DevBuf<uint32_t> cell(n), ccount(mpad), csorted(n), order(n), ckey(n);
        DevBuf<float> dist2(n);
        DevBuf<double> cell_d2(mpad);
        MRS_HIP(hipMemsetAsync(ccount.p, 0, ccount.bytes(), s));
        MRS_HIP(hipMemsetAsync(cell_d2.p, 0, cell_d2.bytes(), s));
        const void* pp = Pp;
        const dim3 ga(div_up(n, 128));
        if (bf16) {
            hipLaunchKernelGGL(k_seed_rows<true>, dim3(mpad), dim3(64), 0, s, pp, KP, KS, n, m,
                               seeds.p, snegh.p);
            hipLaunchKernelGGL(k_assign_cells<true>, ga, dim3(256), 0, s, pp, KP, KS, n, seeds.p,
                               snegh.p, mpad, cell.p, ccount.p, neghp, dist2.p, cell_d2.p);
        } else {
            hipLaunchKernelGGL(k_seed_rows<false>, dim3(mpad), dim3(64), 0, s, pp, KP, KS, n, m,
                               seeds.p, snegh.p);
            hipLaunchKernelGGL(k_assign_cells<false>, ga, dim3(256), 0, s, pp, KP, KS, n, seeds.p,
                               snegh.p, mpad, cell.p, ccount.p, neghp, dist2.p, cell_d2.p);
        }
        MRS_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_cell_keys, dim3(div_up(n, 256)), dim3(256), 0, s, cell.p, dist2.p,
                           cell_d2.p, ccount.p, n, ckey.p);
        MRS_CHECK_LAUNCH();
        radix_sort_pairs_u32(ckey.p, nullptr, n, (m > 1 ? 32 - __builtin_clz(m - 1) : 1) + 1,
                             csorted.p, order.p, s);
        std::vector<uint32_t> hc(m), start(m + 1, 0), pstart(m + 1, 0);
        MRS_HIP(hipMemcpyAsync(hc.data(), ccount.p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, s));
        MRS_HIP(hipStreamSynchronize(s));
        uint64_t pairs1 = 0;
        for (uint32_t c = 0; c < m; c++) {
            const uint32_t padded = (hc[c] + kBlockRows - 1) / kBlockRows * kBlockRows;
            start[c + 1] = start[c] + hc[c];
            pstart[c + 1] = pstart[c] + padded;
            pairs1 += (uint64_t)(padded / kBlockRows) * (padded / (uint32_t)sh.RT);
        }
        if (start[m] != n) throw Error(MRS_DEVICE_ERROR, "kNN cell counts do not add up");
        const uint32_t npos = (pstart[m] + 255u) / 256u * 256u;
        DevBuf<uint32_t> dstart(m + 1), dpstart(m + 1), rowmap(npos), posmap(n);
        MRS_HIP(hipMemcpyAsync(dstart.p, start.data(), sizeof(uint32_t) * (m + 1), hipMemcpyHostToDevice, s));
        MRS_HIP(hipMemcpyAsync(dpstart.p, pstart.data(), sizeof(uint32_t) * (m + 1), hipMemcpyHostToDevice, s));
        MRS_HIP(hipMemsetAsync(rowmap.p, 0xFF, rowmap.bytes(), s));
        hipLaunchKernelGGL(k_cell_layout, dim3(div_up(n, 256)), dim3(256), 0, s, csorted.p, order.p,
                           n, dstart.p, dpstart.p, rowmap.p, posmap.p);
        MRS_CHECK_LAUNCH();
        const int row_bytes = KP * elem;
        DevBuf<unsigned char> P2((size_t)npos * row_bytes);
        DevBuf<float> negh2(npos);
        hipLaunchKernelGGL(k_gather_rows, dim3(div_up((uint64_t)npos * (row_bytes / 16), 256)),
                           dim3(256), 0, s, Pp, row_bytes, neghp, rowmap.p, npos, P2.p, negh2.p);
        MRS_CHECK_LAUNCH();
        const uint32_t nblk = npos / kBlockRows, nst = npos / (uint32_t)sh.RT;
        DevBuf<float> bcen((size_t)nblk * KP), brad(nblk), bnorm(nblk);
        DevBuf<uint32_t> bcnt(nblk);
        DevBuf<float> scen_own, srad_own, snorm_own;
        DevBuf<uint32_t> scnt_own;
        auto bounds = [&](int G, uint32_t ng, float* cen, float* rad, float* gn, uint32_t* cnt) {
            if (bf16)
                hipLaunchKernelGGL(k_group_bounds<true>, dim3(div_up(ng, 4)), dim3(256), 0, s,
                                   (const void*)P2.p, KP, rowmap.p, G, ng, cen, rad, gn, cnt);
            else
                hipLaunchKernelGGL(k_group_bounds<false>, dim3(div_up(ng, 4)), dim3(256), 0, s,
                                   (const void*)P2.p, KP, rowmap.p, G, ng, cen, rad, gn, cnt);
            MRS_CHECK_LAUNCH();
        };
        bounds((int)kBlockRows, nblk, bcen.p, brad.p, bnorm.p, bcnt.p);
        const float *scen = bcen.p, *srad = brad.p, *snorm = bnorm.p;
        const uint32_t* scnt = bcnt.p;
        if ((uint32_t)sh.RT != kBlockRows) {
            scen_own.alloc((size_t)nst * KP);
            srad_own.alloc(nst);
            snorm_own.alloc(nst);
            scnt_own.alloc(nst);
            bounds(sh.RT, nst, scen_own.p, srad_own.p, snorm_own.p, scnt_own.p);
            scen = scen_own.p;
            srad = srad_own.p;
            snorm = snorm_own.p;
            scnt = scnt_own.p;
        }
        DevBuf<uint2> range(nblk);
        hipLaunchKernelGGL(k_block_ranges, dim3(div_up(nblk, 256)), dim3(256), 0, s, dpstart.p, m,
                           nblk, (uint32_t)sh.RT, range.p);
        MRS_CHECK_LAUNCH();
        DevBuf<uint64_t> lists((size_t)npos * S);
        DevBuf<uint64_t> queue((size_t)npos * sh.QC);
        ShortlistArgs sa{};
        sa.P = P2.p;
        sa.negh = negh2.p;
        sa.lists = lists.p;
        sa.queue = queue.p;
        sa.n = n;
        sa.npad = npos;
        sa.KP = KP;
        sa.db_per_split = npos;
        sa.rowmap = rowmap.p;
        // Work items of a pass: (block, list offset, count, aux slot).  A block
        // whose list is longer than chunk stages is split over several items
        // (the first keeps the block's lists, the others fill aux lists that
        // k_merge_aux folds in), and items start longest first, so the last
        // workgroups to start are short ones.
        std::vector<uint32_t> hcnt(nblk), hoff(nblk);
        auto fetch_counts = [&](const uint32_t* dcounts) {
            MRS_HIP(hipMemcpyAsync(hcnt.data(), dcounts, sizeof(uint32_t) * nblk,
                                   hipMemcpyDeviceToHost, s));
            MRS_HIP(hipStreamSynchronize(s));
            uint32_t acc = 0;  // the exclusive scan the device made of the same counts
            for (uint32_t b = 0; b < nblk; b++) {
                hoff[b] = acc;
                acc += hcnt[b];
            }
        };
        std::vector<uint4> hitems, hmerges;
        uint32_t n_aux = 0;
        auto make_items = [&](uint32_t chunk) {
            hitems.clear();
            hmerges.clear();
            n_aux = 0;
            for (uint32_t b = 0; b < nblk; b++) {
                const uint32_t c = hcnt[b];
                if (c == 0) continue;
                const uint32_t parts = (uint32_t)(((uint64_t)c + chunk - 1) / chunk);
                if (parts > 1) hmerges.push_back(make_uint4(b, n_aux, parts - 1, 0u));
                for (uint32_t j = 0; j < parts; j++) {
                    const uint32_t len = std::min(chunk, c - j * chunk);
                    hitems.push_back(make_uint4(b, hoff[b] + j * chunk, len, j ? n_aux++ : kMainList));
                }
            }
            std::stable_sort(hitems.begin(), hitems.end(),
                             [](const uint4& x_, const uint4& y_) { return x_.z > y_.z; });
        };
        DevBuf<uint4> items1, items2, dmerges;
        auto upload = [&](DevBuf<uint4>& dst, const std::vector<uint4>& src) {
            dst.alloc(src.size() ? src.size() : 1);
            if (!src.empty())
                MRS_HIP(hipMemcpyAsync(dst.p, src.data(), sizeof(uint4) * src.size(),
                                       hipMemcpyHostToDevice, s));
        };
        // pass 1: each query block's own cell, as an explicit stage list, one item per block
        DevBuf<uint32_t> counts1(nblk), offs1((size_t)nblk + 1), ids1(pairs1 ? pairs1 : 1);
        hipLaunchKernelGGL(k_range_counts, dim3(div_up(nblk, 256)), dim3(256), 0, s, range.p, nblk,
                           counts1.p);
        MRS_CHECK_LAUNCH();
        exclusive_scan_u32(counts1.p, offs1.p, nblk, offs1.p + nblk, s);
        hipLaunchKernelGGL(k_range_ids, dim3(nblk), dim3(256), 0, s, range.p, offs1.p, ids1.p);
        MRS_CHECK_LAUNCH();
        fetch_counts(counts1.p);
        make_items(0xFFFFFFFFu);
        upload(items1, hitems);
        // pass-1 floors (BF16 kernels whose tiles are one MFMA chain each)
        DevBuf<float> floor1;
        const int ksteps1 = bf16 ? KP / 16 : 0;
        if (bf16 && sh.RT >= 64 && ksteps1 >= 1 && ksteps1 <= 8 && !hitems.empty()) {
            floor1.alloc(npos);
            auto fl = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3((unsigned)hitems.size()), dim3(256), 0, s,
                                   (const u16*)P2.p, negh2.p, range.p, brad.p, items1.p, KP,
                                   sh.RT, S, floor1.p);
            };
            switch (ksteps1) {
                case 1: fl(k_cell_floor<1>); break;
                case 2: fl(k_cell_floor<2>); break;
                case 3: fl(k_cell_floor<3>); break;
                case 4: fl(k_cell_floor<4>); break;
                case 5: fl(k_cell_floor<5>); break;
                case 6: fl(k_cell_floor<6>); break;
                case 7: fl(k_cell_floor<7>); break;
                default: fl(k_cell_floor<8>); break;
            }
            MRS_CHECK_LAUNCH();
            sa.floor = floor1.p;
        }
        sa.items = items1.p;
        sa.stage_ids = ids1.p;
        MRS_HIP(hipStreamSynchronize(s));  // hitems is read by the copy
        mark(1);
        if (!hitems.empty()) run_shortlist(sa, dim3((unsigned)hitems.size(), 1), true);
        // pass 2: the stages the bounds cannot rule out
        DevBuf<uint32_t> counts(nblk), offs((size_t)nblk + 1);
        const dim3 gl(div_up(nblk, kListBlocks));
        hipLaunchKernelGGL(k_stage_lists<false>, gl, dim3(256), 0, s, bcen.p, brad.p, bnorm.p,
                           bcnt.p, scen, srad, snorm, scnt, range.p, negh2.p, lists.p, S, KP, nblk,
                           nst, prune ? 0 : 1, counts.p, nullptr, nullptr);
        MRS_CHECK_LAUNCH();
        exclusive_scan_u32(counts.p, offs.p, nblk, offs.p + nblk, s);
        fetch_counts(counts.p);
        const uint32_t total = nblk ? hoff[nblk - 1] + hcnt[nblk - 1] : 0;
        DevBuf<uint32_t> ids(total ? total : 1);
        hipLaunchKernelGGL(k_stage_lists<true>, gl, dim3(256), 0, s, bcen.p, brad.p, bnorm.p,
                           bcnt.p, scen, srad, snorm, scnt, range.p, negh2.p, lists.p, S, KP, nblk,
                           nst, prune ? 0 : 1, counts.p, offs.p, ids.p);
        MRS_CHECK_LAUNCH();
        // chunks: a tenth of a round of workgroups' share, at least 1024 stages
        // (or the context's chunk, which the tests lower to split small problems)
        const uint32_t chunk = ctx->knn_chunk ? ctx->knn_chunk
                                              : std::max<uint32_t>(1024u, (uint32_t)((uint64_t)total / 2048u));
        make_items(chunk);
        aux_items = n_aux;
        upload(items2, hitems);
        upload(dmerges, hmerges);
        DevBuf<uint64_t> aux_lists((size_t)std::max<uint32_t>(n_aux, 1u) * kBlockRows * S);
        DevBuf<uint64_t> aux_queue((size_t)std::max<uint32_t>(n_aux, 1u) * kBlockRows * sh.QC);
        DevBuf<float> floor(npos);
        hipLaunchKernelGGL(k_list_floor, dim3(div_up(npos, 256)), dim3(256), 0, s, lists.p, S,
                           npos, floor.p);
        MRS_CHECK_LAUNCH();
        MRS_HIP(hipStreamSynchronize(s));  // hitems and hmerges are read by the copies
        sa.items = items2.p;
        sa.stage_ids = ids.p;
        sa.cont = 1;
        sa.aux_lists = aux_lists.p;
        sa.aux_queue = aux_queue.p;
        sa.floor = floor.p;
        if (!hitems.empty()) run_shortlist(sa, dim3((unsigned)hitems.size(), 1), true);
        if (!hmerges.empty()) {
            const dim3 gmg(div_up((uint64_t)hmerges.size() * kBlockRows, 4));
            if (E == 1)
                hipLaunchKernelGGL(k_merge_aux<1>, gmg, dim3(256), 0, s, lists.p, aux_lists.p,
                                   dmerges.p, (uint32_t)hmerges.size());
            else
                hipLaunchKernelGGL(k_merge_aux<2>, gmg, dim3(256), 0, s, lists.p, aux_lists.p,
                                   dmerges.p, (uint32_t)hmerges.size());
            MRS_CHECK_LAUNCH();
        }
        mark(2);
        // the re-rank in cell order, from X in cell order
        const int dp = (d + 3) / 4 * 4;
        DevBuf<float> xs((size_t)npos * dp);
        hipLaunchKernelGGL(k_gather_x, dim3(div_up((uint64_t)npos * dp, 256)), dim3(256), 0, s, x,
                           d, dp, rowmap.p, npos, xs.p);
        const dim3 gr(div_up(npos, 4));
        if (E == 1)
            hipLaunchKernelGGL(k_rerank_cells<1>, gr, dim3(256), 0, s, xs.p, dp, d, npos, lists.p,
                               rowmap.p, posmap.p, k, out_idx, out_dist);
        else
            hipLaunchKernelGGL(k_rerank_cells<2>, gr, dim3(256), 0, s, xs.p, dp, d, npos, lists.p,
                               rowmap.p, posmap.p, k, out_idx, out_dist);
        MRS_CHECK_LAUNCH();
        mark(3);
        MRS_HIP(hipEventSynchronize(ev[3]));  // before the buffers go
        const double full = (double)(pstart[m] / kBlockRows) * (double)(pstart[m] / (uint32_t)sh.RT);
        swept = full > 0 ? ((double)pairs1 + (double)total) / full : 1.0;
        list_max = hcnt.empty() ? 0 : *std::max_element(hcnt.begin(), hcnt.end());
        list_mean = nblk ? (double)total / nblk : 0.0;
    }
    if (timings) {
        float a = 0, b = 0, c = 0;
        MRS_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        MRS_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
        MRS_HIP(hipEventElapsedTime(&c, ev[2], ev[3]));
        timings->pack_ms = a;
        timings->shortlist_ms = b;
        timings->rerank_ms = c;
        timings->splits = splits;
        timings->shortlist = S;
        timings->swept = swept;
        timings->cells = (int)cells;
        timings->list_max = list_max;
        timings->list_mean = list_mean;
        timings->aux_items = aux_items;
    }
}

// ------------------------------------------- the search without allocation
//
// Scratch for mrs_knn_exhaustive_noalloc: the full sweep's packed operand,
// norms, candidate lists and queues for up to kNoAllocRows rows of up to 128
// dimensions, shortlists of 64 and four database splits (72 MiB).  One copy
// lives on every GPU the library runs on, until the process ends.  Host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (the array's name in host code is not a device address).
constexpr int64_t kNoAllocRows = 16384;
constexpr int kNoAllocSplits = 4;
constexpr size_t kNoAllocScratchBytes =
    (size_t)kNoAllocRows * 128 * 4                                  // packed operand (FP32)
    + (size_t)kNoAllocRows * 4                                      // -|y|^2 / 2
    + 2 * (size_t)kNoAllocSplits * kNoAllocRows * 64 * 8            // lists and queues
    + 4 * 256;                                                      // alignment of the pieces

namespace {

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_addr;  // device -> address of its scratch

// the current device's scratch address, looked up once per device
unsigned char* device_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) return it->second;
    void* ptr = nullptr;
    MRS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(MRS_DEVICE_ERROR, "no address for the kNN scratch");
    g_scratch_addr.emplace(device, static_cast<unsigned char*>(ptr));
    return static_cast<unsigned char*>(ptr);
}

}  // namespace

}  // namespace mrs

extern "C" {

int mrs_knn_exhaustive_noalloc(mrs_ctx* ctx, const float* x, int64_t n, int32_t d, int32_t k,
                               int32_t precision, int32_t shortlist, int32_t* idx, float* dist) {
    return mrs::guarded([&] {
        if (!ctx || !x || !idx || !dist) throw mrs::Error(MRS_INVALID_ARGUMENT, "NULL argument");
        if (n > mrs::kNoAllocRows)
            throw mrs::Error(MRS_INVALID_ARGUMENT,
                             "n = " + std::to_string(n) + " exceeds the no-allocation search's " +
                                 std::to_string(mrs::kNoAllocRows) + " rows");
        if (std::max(shortlist, k) > 64)
            throw mrs::Error(MRS_INVALID_ARGUMENT,
                             "the no-allocation search keeps shortlists of at most 64");
        MRS_HIP(hipSetDevice(ctx->device));
        if (!mrs::is_device_ptr(x) || !mrs::is_device_ptr(idx) || !mrs::is_device_ptr(dist))
            throw mrs::Error(MRS_INVALID_ARGUMENT,
                             "the no-allocation search takes device buffers only (staging host "
                             "memory would allocate)");
        mrs::KnnScratch scr{mrs::device_scratch(ctx->device), mrs::kNoAllocScratchBytes,
                            mrs::kNoAllocSplits};
        mrs::StreamScope scope(ctx->stream);
        mrs::knn_exhaustive_dev(ctx, x, n, d, k, precision, shortlist, idx, dist, nullptr, &scr);
    });
}

int mrs_knn_exhaustive(mrs_ctx* ctx, const float* x, int64_t n, int32_t d, int32_t k,
                       int32_t precision, int32_t shortlist, int32_t* idx, float* dist) {
    return mrs::guarded([&] {
        if (!ctx || !x || !idx || !dist) throw mrs::Error(MRS_INVALID_ARGUMENT, "NULL argument");
        MRS_HIP(hipSetDevice(ctx->device));
        mrs::StreamScope scope(ctx->stream, ctx->pool);
        mrs::DevView<float> xv(x, (size_t)n * d, ctx->stream);
        mrs::DevView<int32_t> iv(idx, (size_t)n * k, ctx->stream, false);
        mrs::DevView<float> dv(dist, (size_t)n * k, ctx->stream, false);
        mrs::knn_exhaustive_dev(ctx, xv.p, n, d, k, precision, shortlist, iv.p, dv.p, nullptr);
        iv.download(ctx->stream);
        dv.download(ctx->stream);
        MRS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int mrs_knn_exhaustive_timed(mrs_ctx* ctx, const float* x, int64_t n, int32_t d, int32_t k,
                             int32_t precision, int32_t shortlist, int32_t* idx, float* dist,
                             double* ms_out) {
    return mrs::guarded([&] {
        if (!ctx || !x || !idx || !dist) throw mrs::Error(MRS_INVALID_ARGUMENT, "NULL argument");
        MRS_HIP(hipSetDevice(ctx->device));
        mrs::StreamScope scope(ctx->stream, ctx->pool);
        mrs::DevView<float> xv(x, (size_t)n * d, ctx->stream);
        mrs::DevView<int32_t> iv(idx, (size_t)n * k, ctx->stream, false);
        mrs::DevView<float> dv(dist, (size_t)n * k, ctx->stream, false);
        mrs::KnnTimings t;
        mrs::knn_exhaustive_dev(ctx, xv.p, n, d, k, precision, shortlist, iv.p, dv.p, &t);
        iv.download(ctx->stream);
        dv.download(ctx->stream);
        MRS_HIP(hipStreamSynchronize(ctx->stream));
        if (ms_out) {
            ms_out[0] = t.pack_ms;
            ms_out[1] = t.shortlist_ms;
            ms_out[2] = t.rerank_ms;
            ms_out[3] = t.splits;
            ms_out[4] = t.shortlist;
            ms_out[5] = t.swept;
            ms_out[6] = t.cells;
            ms_out[7] = t.list_max;
            ms_out[8] = t.list_mean;
            ms_out[9] = t.aux_items;
        }
    });
}

}  // extern "C"
