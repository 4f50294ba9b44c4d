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
//   Each workgroup holds 8 waves x 32 queries; database tiles of 64 rows are
//   staged through LDS (rows padded by 16 B against bank conflicts) and shared
//   by the 8 waves.  Candidates beating the query's threshold go to a per-query
//   LDS queue (64 entries); a queue that could overflow on the next tile is
//   merged into the query's sorted top-S list in global memory by the whole
//   wavefront (bitonic sort + merge on 64-bit keys), which also refreshes the
//   threshold.  Large problems use one database split; small ones split the
//   database so the grid still fills 256 CUs, and a merge kernel joins the
//   split lists.
//
// Stage 2 -- exact re-rank: one wavefront per query recomputes the S candidate
//   distances in FP32 as fmaf((x_q[t]-x_c[t])^2) over t in order (the same
//   arithmetic as the CPU oracle) and sorts (distance, index) keys; the first k
//   are the result.  This makes the indices exact whenever the true top-k is in
//   the shortlist, for both precisions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "knn.hpp"

namespace mrs {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef unsigned short u16;

constexpr int kWaves = 8;          // waves per workgroup in the shortlist kernel
constexpr int kQPerWave = 32;      // queries per wave (MFMA B columns)
constexpr int kTileRows = 64;      // database rows per LDS tile (two MFMA A tiles)
constexpr int kQCap = 64;          // per-query queue capacity
constexpr int kPF = 4;             // database tiles in flight (register prefetch)
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

__device__ inline uint64_t shfl_xor64(uint64_t v, int m) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor(lo, m, 64);
    hi = __shfl_xor(hi, m, 64);
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
    uint32_t n, npad;
    int KP;
    uint32_t db_per_split;  // multiple of kTileRows
};

template <bool BF16, int E, int CPT>
__global__ __launch_bounds__(kWaves * 64) void k_shortlist(ShortlistArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int S = 64 * E;
    const int KP = a.KP;
    const int elem = BF16 ? 2 : 4;
    const int row_bytes = KP * elem + 16;  // padded LDS row
    unsigned char* tile = smem;                                      // [kTileRows][row_bytes]
    float* tile_negh = reinterpret_cast<float*>(smem + kTileRows * row_bytes);  // [64]
    uint64_t* queue = reinterpret_cast<uint64_t*>(tile_negh + kTileRows);      // [waves][32][QCap]
    uint32_t* qcnt = reinterpret_cast<uint32_t*>(queue + kWaves * kQPerWave * kQCap);
    float* qthr = reinterpret_cast<float*>(qcnt + kWaves * kQPerWave);

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, h = lane >> 5, col = lane & 31;
    const uint32_t qbase = (blockIdx.x * kWaves + wave) * kQPerWave;
    const uint32_t split = blockIdx.y;
    const uint32_t db0 = split * a.db_per_split;
    const uint32_t db1 = min(db0 + a.db_per_split, a.npad);
    uint64_t* lists = a.lists + ((size_t)split * a.npad + qbase) * S;
    uint64_t* myq = queue + (size_t)wave * kQPerWave * kQCap;
    uint32_t* mycnt = qcnt + wave * kQPerWave;
    float* mythr = qthr + wave * kQPerWave;

    // init this wave's lists, counters and thresholds
    for (int i = lane; i < kQPerWave * S; i += 64) lists[i] = kInfKey;
    if (lane < kQPerWave) {
        mycnt[lane] = 0;
        mythr[lane] = -INFINITY;
    }
    // B fragments: this lane's query (column) over its half of K
    const uint32_t q = qbase + col;
    constexpr int MAXSTEPS = BF16 ? 8 : 64;  // KP <= 128
    const int steps = BF16 ? KP / 16 : KP / 2;
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
    __syncthreads();

    // wave-wide merge of query b's queue into its list
    auto merge_query = [&](int b) {
        const uint32_t c = mycnt[b];
        uint64_t qv = (uint32_t)lane < c ? myq[b * kQCap + lane] : kInfKey;
        uint64_t one[1] = {qv};
        wave_sort<1>(one);
        uint64_t v[E];
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = lists[(size_t)b * S + e * 64 + lane];
        wave_merge<E>(v, one[0]);
#pragma unroll
        for (int e = 0; e < E; e++) lists[(size_t)b * S + e * 64 + lane] = v[e];
        const uint64_t last = __shfl((uint32_t)(v[E - 1] >> 32), 63, 64);
        const uint64_t lastlo = __shfl((uint32_t)v[E - 1], 63, 64);
        if (lane == 0) {
            mythr[b] = key_score((last << 32) | lastlo);
            mycnt[b] = 0;
        }
    };

    // epilogue of one 32-row MFMA tile: any element beating the query's threshold?
    auto epilogue = [&](const f32x16& acc, uint32_t rowbase) {
        float m = acc[0];
#pragma unroll
        for (int r = 1; r < 16; r++) m = fmaxf(m, acc[r]);
        if (__ballot(m > thr) == 0ull) return;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint32_t db = rowbase + (r >> 2) * 8 + h * 4 + (r & 3);
            if (acc[r] > thr && db < a.n && db != q) {
                const uint32_t pos = atomicAdd(&mycnt[col], 1u);
                myq[col * kQCap + pos] = score_key(acc[r], db);
            }
        }
        // flush queues that the next 32-row tile could overflow
        const bool full = lane < kQPerWave && mycnt[lane] > (uint32_t)(kQCap - 32);
        uint64_t need = __ballot(full);
        while (need) {
            const int b = __builtin_ctzll(need);
            need &= need - 1;
            merge_query(b);
        }
        thr = mythr[col];
    };

    // both 32-row MFMA tiles of the staged 64-row tile, as two independent chains
    auto process_tile = [&](uint32_t t0) {
        f32x16 acc0, acc1;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float4 n0 = *reinterpret_cast<const float4*>(tile_negh + i * 8 + h * 4);
            const float4 n1 = *reinterpret_cast<const float4*>(tile_negh + 32 + i * 8 + h * 4);
            acc0[i * 4 + 0] = n0.x; acc0[i * 4 + 1] = n0.y; acc0[i * 4 + 2] = n0.z; acc0[i * 4 + 3] = n0.w;
            acc1[i * 4 + 0] = n1.x; acc1[i * 4 + 1] = n1.y; acc1[i * 4 + 2] = n1.z; acc1[i * 4 + 3] = n1.w;
        }
        if (BF16) {
            const bf16x8* a0 = reinterpret_cast<const bf16x8*>(tile + col * row_bytes + h * KP);
            const bf16x8* a1 = reinterpret_cast<const bf16x8*>(tile + (32 + col) * row_bytes + h * KP);
#pragma unroll
            for (int st = 0; st < MAXSTEPS; st++)
                if (st < steps) {
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0[st], bq[st], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[st], bq[st], acc1, 0, 0, 0);
                }
        } else {
            const float* a0 = reinterpret_cast<const float*>(tile + col * row_bytes) + h * (KP / 2);
            const float* a1 = reinterpret_cast<const float*>(tile + (32 + col) * row_bytes) + h * (KP / 2);
#pragma unroll
            for (int st = 0; st < MAXSTEPS; st++)
                if (st < steps) {
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[st], bqf[st], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1[st], bqf[st], acc1, 0, 0, 0);
                }
        }
        epilogue(acc0, t0);
        epilogue(acc1, t0 + 32);
    };

    // Database tiles stream through registers kPF tiles ahead of the LDS stage,
    // so the global/L2 latency of tile t+kPF hides behind the MFMAs of tile t.
    const int cpr = KP * elem / 16;           // 16-byte chunks per row
    const int tile_chunks = kTileRows * cpr;  // contiguous in P
    const uint32_t ntiles = (db1 - db0) / kTileRows;
    uint4 pre[kPF][CPT];
    float pre_n[kPF];
    auto load_tile = [&](uint4 (&dst)[CPT], float& dn, uint32_t t) {
        const uint32_t t0 = db0 + t * kTileRows;
        const uint4* base = reinterpret_cast<const uint4*>(
            reinterpret_cast<const unsigned char*>(a.P) + (size_t)t0 * KP * elem);
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * kWaves * 64;
            if (c < tile_chunks) dst[j] = base[c];
        }
        if (tid < kTileRows) dn = a.negh[t0 + tid];
    };
    auto store_tile = [&](const uint4 (&src)[CPT], float dn) {
#pragma unroll
        for (int j = 0; j < CPT; j++) {
            const int c = tid + j * kWaves * 64;
            if (c < tile_chunks) {
                const int r = c / cpr, o = c - r * cpr;
                *reinterpret_cast<uint4*>(tile + r * row_bytes + o * 16) = src[j];
            }
        }
        if (tid < kTileRows) tile_negh[tid] = dn;
    };
#pragma unroll
    for (int p = 0; p < kPF; p++)
        if ((uint32_t)p < ntiles) load_tile(pre[p], pre_n[p], p);
    for (uint32_t tb = 0; tb < ntiles; tb += kPF) {
#pragma unroll
        for (int p = 0; p < kPF; p++) {
            const uint32_t t = tb + p;
            if (t < ntiles) {
                __syncthreads();  // the previous tile's LDS reads are done
                store_tile(pre[p], pre_n[p]);
                if (t + kPF < ntiles) load_tile(pre[p], pre_n[p], t + kPF);
                __syncthreads();
                process_tile(db0 + t * kTileRows);
            }
        }
    }
    // final flush of every non-empty queue
    {
        const bool any = lane < kQPerWave && mycnt[lane] > 0;
        uint64_t need = __ballot(any);
        while (need) {
            const int b = __builtin_ctzll(need);
            need &= need - 1;
            merge_query(b);
        }
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

// exact FP32 re-rank of the shortlist; one wave per query
template <int E>
__global__ __launch_bounds__(256) void k_rerank(const float* __restrict__ x, uint32_t n, int d,
                                                const uint64_t* __restrict__ lists, int k,
                                                int32_t* __restrict__ out_idx,
                                                float* __restrict__ out_dist) {
    constexpr int S = 64 * E;
    const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= n) return;
    const float* xq = x + (size_t)q * d;
    uint64_t key[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const uint64_t lk = lists[(size_t)q * S + e * 64 + lane];
        if (lk == kInfKey) {
            key[e] = kInfKey;
            continue;
        }
        const uint32_t c = (uint32_t)lk;
        const float* xc = x + (size_t)c * d;
        float s = 0.0f;
        for (int t = 0; t < d; t++) {
            const float df = xq[t] - xc[t];
            s = fmaf(df, df, s);
        }
        key[e] = ((uint64_t)__float_as_uint(s) << 32) | c;
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

size_t shortlist_lds_bytes(bool bf16, int KP) {
    const int elem = bf16 ? 2 : 4;
    const size_t row_bytes = (size_t)KP * elem + 16;
    return kTileRows * row_bytes + kTileRows * sizeof(float) +
           (size_t)kWaves * kQPerWave * kQCap * sizeof(uint64_t) +
           (size_t)kWaves * kQPerWave * (sizeof(uint32_t) + sizeof(float));
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
    const int KP = bf16 ? ((d + 15) / 16) * 16 : ((d + 3) / 4) * 4;
    const uint32_t qpb = kWaves * kQPerWave;  // queries per block
    const uint32_t npad = ((n + qpb - 1) / qpb) * qpb;  // also a multiple of kTileRows
    hipStream_t s = ctx->stream;

    hipEvent_t ev[4];
    for (auto& e : ev) MRS_HIP(hipEventCreate(&e));
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

    const uint32_t qblocks = npad / qpb;
    int splits = (int)std::min<uint32_t>(8, std::max<uint32_t>(1, (1024 + qblocks - 1) / qblocks));
    uint32_t db_per = ((npad + splits - 1) / splits + kTileRows - 1) / kTileRows * kTileRows;
    splits = (int)((npad + db_per - 1) / db_per);
    const int S = 64 * E;
    DevBuf<uint64_t> lists((size_t)splits * npad * S);

    ShortlistArgs sa;
    sa.P = P.p;
    sa.negh = negh.p;
    sa.lists = lists.p;
    sa.n = n;
    sa.npad = npad;
    sa.KP = KP;
    sa.db_per_split = db_per;
    const size_t lds = shortlist_lds_bytes(bf16, KP);
    if (lds > 160 * 1024) throw Error(MRS_INVALID_ARGUMENT, "d too large for the shortlist kernel");
    auto launch = [&](auto kern) {
        MRS_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
        hipLaunchKernelGGL(kern, dim3(qblocks, splits), dim3(kWaves * 64), lds, s, sa);
    };
    MRS_HIP(hipEventRecord(ev[1], s));
    // 16-byte chunks each thread stages per 64-row tile
    const int cpt = (kTileRows * (KP * (bf16 ? 2 : 4) / 16) + kWaves * 64 - 1) / (kWaves * 64);
    if (bf16) {
        if (cpt == 1) E == 1 ? launch(k_shortlist<true, 1, 1>) : launch(k_shortlist<true, 2, 1>);
        else E == 1 ? launch(k_shortlist<true, 1, 2>) : launch(k_shortlist<true, 2, 2>);
    } else {
        if (cpt == 1) E == 1 ? launch(k_shortlist<false, 1, 1>) : launch(k_shortlist<false, 2, 1>);
        else if (cpt == 2) E == 1 ? launch(k_shortlist<false, 1, 2>) : launch(k_shortlist<false, 2, 2>);
        else E == 1 ? launch(k_shortlist<false, 1, 4>) : launch(k_shortlist<false, 2, 4>);
    }
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
    if (E == 1)
        hipLaunchKernelGGL(k_rerank<1>, dim3(div_up(n, 4)), dim3(256), 0, s, x, n, d, lists.p, k,
                           out_idx, out_dist);
    else
        hipLaunchKernelGGL(k_rerank<2>, dim3(div_up(n, 4)), dim3(256), 0, s, x, n, d, lists.p, k,
                           out_idx, out_dist);
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
    for (auto& e : ev) (void)hipEventDestroy(e);
}

}  // namespace mrs

extern "C" {

int mrs_knn_exhaustive(mrs_ctx* ctx, const float* x, int64_t n, int32_t d, int32_t k,
                       int32_t precision, int32_t shortlist, int32_t* idx, float* dist) {
    return mrs::guarded([&] {
        if (!ctx || !x || !idx || !dist) throw mrs::Error(MRS_INVALID_ARGUMENT, "NULL argument");
        MRS_HIP(hipSetDevice(ctx->device));
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
