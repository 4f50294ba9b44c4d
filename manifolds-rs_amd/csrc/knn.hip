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
//   dropped at insertion.  Large problems use one database split; small ones
//   split the database so the grid still fills 256 CUs, and a merge kernel
//   joins the split lists.
//
// Stage 2 -- exact re-rank: one wavefront per query recomputes the S candidate
//   distances in FP32 as fmaf((x_q[t]-x_c[t])^2) over t in order (the same
//   arithmetic as the CPU oracle) and sorts (distance, index) keys; the first k
//   are the result.  This makes the indices exact whenever the true top-k is in
//   the shortlist, for both precisions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cfloat>
#include <cmath>

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
};

// Wave-wide merge of query b's queue (c entries) into its sorted global list;
// returns the list's new threshold (the score of its last entry).
template <int E, int QC>
__device__ inline float merge_query(int b, uint32_t c, int lane, uint64_t* __restrict__ lists,
                                    const uint64_t* __restrict__ myq) {
    constexpr int S = 64 * E;
    // the queue entries were stored by this wave's lanes: complete those
    // stores before any lane reads another lane's entry
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    uint64_t one[1] = {(uint32_t)lane < c ? myq[b * QC + lane] : kInfKey};
    wave_sort<1>(one);
    uint64_t v[E];
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = lists[(size_t)b * S + e * 64 + lane];
    wave_merge<E>(v, one[0]);
#pragma unroll
    for (int e = 0; e < E; e++) lists[(size_t)b * S + e * 64 + lane] = v[e];
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v[E - 1] >> 32), 63);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v[E - 1], 63);
    return key_score(((uint64_t)hi << 32) | lo);
}

// Epilogue of one 32-row MFMA tile: the lane's 16 scores (one query, rows
// rowbase + (r/4)*8 + h*4 + r%4) against the query's threshold.  A query's
// queue count qn and threshold thr live in registers, identical in its two
// lanes (h = 0, 1); the halves learn each other's insert count with one lane
// swap, so queueing costs no LDS round trip.  Most tiles leave on the
// max/ballot test; a queue that could overflow on the next tile is merged.
template <int E, int QC>
__device__ inline void epilogue(f32x16 acc, uint32_t rowbase, uint32_t q, int h, int col,
                                int lane, float& thr, uint32_t& qn, uint64_t* __restrict__ lists,
                                uint64_t* __restrict__ myq) {
    float gm[4];  // maxima of the four 4-row groups (rows i*8 + h*4 + 0..3)
#pragma unroll
    for (int g = 0; g < 4; g++)
        gm[g] = fmaxf(fmaxf(acc[g * 4], acc[g * 4 + 1]), fmaxf(acc[g * 4 + 2], acc[g * 4 + 3]));
    const float m = fmaxf(fmaxf(gm[0], gm[1]), fmaxf(gm[2], gm[3]));
    if (__ballot(m > thr) == 0ull) return;
    // The query's own row scores highest and may pass the test above; it is
    // never queued.  Excluding it here, on the rare path, keeps the common
    // path free of per-element index compares.
    const uint32_t self = q - rowbase - (uint32_t)h * 4u;  // q's offset among this lane's rows
    auto take = [&](int r) {
        return acc[r] > thr && (uint32_t)((r >> 2) * 8 + (r & 3)) != self;
    };
    uint32_t c = 0;
#pragma unroll
    for (int r = 0; r < 16; r++) c += take(r) ? 1u : 0u;
    const uint32_t co = __shfl_xor(c, 32, 64);
    uint32_t pos = qn + (h ? co : 0u);
    qn += c + co;
    uint64_t* dst = myq + col * QC;
    // a group no lane of the wave passes is skipped as a whole
#pragma unroll
    for (int g = 0; g < 4; g++)
        if (gm[g] > thr) {
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (take(g * 4 + j))
                    dst[pos++] = score_key(acc[g * 4 + j], rowbase + g * 8 + h * 4 + j);
        }
    uint64_t need = __ballot(qn > (uint32_t)(QC - 32)) & 0xFFFFFFFFull;  // lanes 0..31: one per query
    while (need) {
        const int b = __builtin_ctzll(need);
        need &= need - 1;
        const float t = merge_query<E, QC>(b, __builtin_amdgcn_readlane(qn, b), lane, lists, myq);
        if (col == b) {
            qn = 0;
            thr = t;
        }
    }
}

// Global -> registers: this thread's 16-byte chunks of an RT-row stage (rows
// are contiguous in P) and, for tid < RT, the row's -0.5|y|^2.
template <int CPT, int RT, int W>
__device__ inline void stage_load(u32x4 (&r)[CPT], float& rn, const unsigned char* __restrict__ P,
                                  const float* __restrict__ negh, uint32_t row0, int row_elems_bytes,
                                  int chunks, int tid) {
    const u32x4* base = reinterpret_cast<const u32x4*>(P + (size_t)row0 * row_elems_bytes);
#pragma unroll
    for (int j = 0; j < CPT; j++) {
        const int c = tid + j * W * 64;
        if (c < chunks) r[j] = base[c];
    }
    if (tid < RT) rn = negh[row0 + tid];
}

// Registers -> LDS stage buffer (rows padded to row_bytes against bank conflicts).
template <int CPT, int RT, int W>
__device__ inline void stage_store(const u32x4 (&r)[CPT], float rn, unsigned char* tile,
                                   float* tile_negh, int row_bytes, int cpr, int chunks, int tid) {
#pragma unroll
    for (int j = 0; j < CPT; j++) {
        const int c = tid + j * W * 64;
        if (c < chunks) {
            const int row = c / cpr, o = c - row * cpr;
            *reinterpret_cast<u32x4*>(tile + row * row_bytes + o * 16) = r[j];
        }
    }
    if (tid < RT) tile_negh[tid] = rn;
}

// The distance GEMM + top-S filter.  One workgroup = W waves x 32 queries; the
// database split [db0, db1) streams through two LDS stage buffers of RT rows
// (double buffering: one barrier per stage), each stage prefetched into
// registers one stage ahead.  QC = per-query queue capacity.  STEPS > 0
// compiles the K-step count in (BF16: KP / 16), so the query fragments take
// only the registers they need and the MFMA chain is straight-line code.
//   The loop is latency-bound (waves mostly wait on LDS, the stage barrier
// and the prefetch), so occupancy pays: the BF16 kernels with the step count
// compiled in are held to 128 registers (four waves per SIMD; at most a few
// spilled dwords, none for the 128-row stages), the FP32 ones to 168 (three).
template <bool BF16, int E, int RT, int QC, int W, int STEPS>
__global__ __launch_bounds__(W * 64)
__attribute__((amdgpu_waves_per_eu(BF16 && STEPS > 0 ? 4 : (BF16 ? 1 : 3)))) void k_shortlist(ShortlistArgs a) {
    constexpr int CPT = (kStageChunks * kWavesMax) / W;
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
    const uint32_t qbase = (blockIdx.x * W + wave) * kQPerWave;
    const uint32_t split = blockIdx.y;
    const uint32_t db0 = split * a.db_per_split;
    const uint32_t db1 = min(db0 + a.db_per_split, a.npad);
    uint64_t* lists = a.lists + ((size_t)split * a.npad + qbase) * S;
    uint64_t* myq = a.queue + ((size_t)split * a.npad + qbase) * QC;

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
    float thr = -INFINITY;
    uint32_t qn = 0;

    const unsigned char* P = reinterpret_cast<const unsigned char*>(a.P);
    const int row_elems_bytes = KP * elem;
    const int cpr = row_elems_bytes / 16;  // 16-byte chunks per row
    const int chunks = RT * cpr;
    const uint32_t nst = (db1 - db0) / RT;
    u32x4 R[CPT];
    float Rn = 0.0f;
    if (nst > 0) {
        stage_load<CPT, RT, W>(R, Rn, P, a.negh, db0, row_elems_bytes, chunks, tid);
        stage_store<CPT, RT, W>(R, Rn, tile0, negh0, row_bytes, cpr, chunks, tid);
    }
    // everything loaded so far (query fragments, stage 0) is complete here, so
    // inside the loop only the one-stage-ahead prefetch is ever in flight and
    // the MFMAs do not wait on it
    __builtin_amdgcn_s_waitcnt(0);
    if (nst > 1) stage_load<CPT, RT, W>(R, Rn, P, a.negh, db0 + RT, row_elems_bytes, chunks, tid);
    __syncthreads();

    for (uint32_t st = 0; st < nst; st++) {
        const bool odd = st & 1u;
        const unsigned char* cur = odd ? tile1 : tile0;
        const float* curn = odd ? negh1 : negh0;
        // stage st+1 -> the other buffer (last read in stage st-1, before the
        // barrier that closed it), then fetch stage st+2 into registers
        if (st + 1 < nst)
            stage_store<CPT, RT, W>(R, Rn, odd ? tile0 : tile1, odd ? negh0 : negh1, row_bytes, cpr,
                                    chunks, tid);
        if (st + 2 < nst)
            stage_load<CPT, RT, W>(R, Rn, P, a.negh, db0 + (st + 2) * RT, row_elems_bytes, chunks,
                                   tid);
        const uint32_t r0 = db0 + st * RT;
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
            epilogue<E, QC>(acc0, r0 + sp * 32, q, h, col, lane, thr, qn, lists, myq);
            if (kPair) epilogue<E, QC>(acc1, r0 + (sp + 1) * 32, q, h, col, lane, thr, qn, lists, myq);
        }
        __syncthreads();
    }
    // final flush of every non-empty queue
    uint64_t need = __ballot(qn > 0) & 0xFFFFFFFFull;
    while (need) {
        const int b = __builtin_ctzll(need);
        need &= need - 1;
        merge_query<E, QC>(b, __builtin_amdgcn_readlane(qn, b), lane, lists, myq);
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
    uint64_t key[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint64_t lk = lists[(size_t)q * S + e * 64 + lane];
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
                        KnnTimings* timings) {
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

    // stage timing events, destroyed on every exit (including a throw)
    struct Events {
        hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Events() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } evs;
    hipEvent_t* ev = evs.e;
    for (int i = 0; i < 4; i++) MRS_HIP(hipEventCreate(&ev[i]));
    MRS_HIP(hipEventRecord(ev[0], s));

    DevBuf<unsigned char> P((size_t)npad * KP * (bf16 ? 2 : 4));
    DevBuf<float> negh(npad);
    if (bf16)
        hipLaunchKernelGGL(k_pack<true>, dim3(div_up(npad, 4)), dim3(256), 0, s, x, n, d, npad, KP,
                           (void*)P.p, negh.p);
    else
        hipLaunchKernelGGL(k_pack<false>, dim3(div_up(npad, 4)), dim3(256), 0, s, x, n, d, npad,
                           KP, (void*)P.p, negh.p);
    MRS_CHECK_LAUNCH();

    const Shape sh = shortlist_shape(bf16, KP);
    const uint32_t qblocks = npad / (uint32_t)(sh.W * kQPerWave);
    int splits = (int)std::min<uint32_t>(8, std::max<uint32_t>(1, (1024 + qblocks - 1) / qblocks));
    uint32_t db_per = ((npad + splits - 1) / splits + kSplitAlign - 1) / kSplitAlign * kSplitAlign;
    splits = (int)((npad + db_per - 1) / db_per);
    const int S = 64 * E;
    DevBuf<uint64_t> lists((size_t)splits * npad * S);
    DevBuf<uint64_t> queue((size_t)splits * npad * sh.QC);

    ShortlistArgs sa;
    sa.P = P.p;
    sa.negh = negh.p;
    sa.lists = lists.p;
    sa.queue = queue.p;
    sa.n = n;
    sa.npad = npad;
    sa.KP = KP;
    sa.db_per_split = db_per;
    const size_t lds = shortlist_lds_bytes(bf16, KP, sh.RT);
    auto launch = [&](auto kern, int W) {
        ensure_dynamic_lds((const void*)kern, lds);
        hipLaunchKernelGGL(kern, dim3(qblocks, splits), dim3(W * 64), lds, s, sa);
    };
    MRS_HIP(hipEventRecord(ev[1], s));
    bool launched = false;
    // BF16 kernels are compiled per K-step count for the shapes that take it
    // (128-row stages: KP <= 64; 64-row stages: KP 80..128)
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
                launch(k_shortlist<B, EE, RTT, QCC, WW, ST>, WW);
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
            launch(k_shortlist<B, EE, RTT, QCC, WW, 0>, WW);
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
    MRS_HIP(hipEventRecord(ev[2], s));
    if (splits > 1) {
        if (E == 1)
            hipLaunchKernelGGL(k_merge_splits<1>, dim3(div_up(npad, 4)), dim3(256), 0, s, lists.p,
                               npad, splits);
        else
            hipLaunchKernelGGL(k_merge_splits<2>, dim3(div_up(npad, 4)), dim3(256), 0, s, lists.p,
                               npad, splits);
        MRS_CHECK_LAUNCH();
    }
    const size_t rlds = rerank_lds_bytes(d);
    auto rerank = [&](auto kern) {
        ensure_dynamic_lds((const void*)kern, rlds);
        hipLaunchKernelGGL(kern, dim3(div_up(n, kRerankWaves)), dim3(kRerankWaves * 64), rlds, s, x,
                           n, d, lists.p, k, out_idx, out_dist);
    };
    if (E == 1) rerank(k_rerank<1>);
    else rerank(k_rerank<2>);
    MRS_CHECK_LAUNCH();
    MRS_HIP(hipEventRecord(ev[3], s));
    MRS_HIP(hipEventSynchronize(ev[3]));
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
    }
}

}  // namespace mrs

extern "C" {

int mrs_knn_exhaustive(mrs_ctx* ctx, const float* x, int64_t n, int32_t d, int32_t k,
                       int32_t precision, int32_t shortlist, int32_t* idx, float* dist) {
    return mrs::guarded([&] {
        if (!ctx || !x || !idx || !dist) throw mrs::Error(MRS_INVALID_ARGUMENT, "NULL argument");
        MRS_HIP(hipSetDevice(ctx->device));
        mrs::StreamScope scope(ctx->stream);
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
        mrs::StreamScope scope(ctx->stream);
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
        }
    });
}

}  // extern "C"
