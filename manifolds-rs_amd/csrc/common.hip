// Context, error plumbing, pointer classification and a device exclusive scan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <map>
#include <mutex>
#include <string>
#include <utility>

#include "epoch_common.hpp"
#include "mrs_internal.hpp"

namespace mrs {

static thread_local std::string g_last_error;
static thread_local hipStream_t t_stream = nullptr;
static thread_local bool t_stream_set = false;

static thread_local std::shared_ptr<DevicePool> t_pool;

StreamScope::StreamScope(hipStream_t s, std::shared_ptr<DevicePool> pool)
    : prev(t_stream), prev_set(t_stream_set), prev_pool(t_pool) {
    t_stream = s;
    t_stream_set = true;
    t_pool = std::move(pool);
}
StreamScope::~StreamScope() {
    t_stream = prev;
    t_stream_set = prev_set;
    t_pool = std::move(prev_pool);
}
bool current_stream(hipStream_t* s) {
    if (t_stream_set) *s = t_stream;
    return t_stream_set;
}
std::shared_ptr<DevicePool> current_pool() { return t_pool; }

// ------------------------------------------------------------ memory pool

HipRT::Event HipRT::record(Stream s) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess) {
        if (hipEventRecord(e, s) == hipSuccess) return e;
        (void)hipEventDestroy(e);
    }
    (void)hipGetLastError();
    (void)hipStreamSynchronize(s);  // no event: the stream's work is done instead
    return nullptr;
}

void HipRT::wait(Stream s, Event e) {
    if (e && hipStreamWaitEvent(s, e, 0) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipEventSynchronize(e);
    }
}

void HipRT::host_wait(Event e) {
    if (e) (void)hipEventSynchronize(e);
}

void HipRT::destroy(Event e) {
    if (e) (void)hipEventDestroy(e);
}

void* HipRT::alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

void HipRT::release(void* p) { (void)hipFree(p); }

void HipRT::sync_all() { (void)hipDeviceSynchronize(); }

DevicePool::~DevicePool() {
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(device_);
    cache_.trim();
    if (prev >= 0) (void)hipSetDevice(prev);
}

void* DevicePool::get(size_t bytes, size_t* got) {
    hipStream_t s = nullptr;
    const bool scoped = current_stream(&s);
    void* p = cache_.get(bytes, got, scoped ? s : nullptr);
    if (!p)
        throw Error(MRS_DEVICE_ERROR, "hipMalloc of " + std::to_string(bytes) +
                                          " bytes failed (device memory exhausted)");
    return p;
}

void DevicePool::put(void* p, size_t bytes) {
    hipStream_t s = nullptr;
    const bool scoped = current_stream(&s);
    cache_.put(p, bytes, scoped ? s : nullptr);
}

void DevicePool::trim() { cache_.trim(); }

void set_last_error(const std::string& msg) { g_last_error = msg; }

bool is_device_ptr(const void* ptr) {
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, ptr);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // unregistered host memory: clear the sticky error
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

namespace {

constexpr int kScanWg = 256;
constexpr int kScanItems = 8;  // per thread
constexpr int kScanTile = kScanWg * kScanItems;

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Block-wide exclusive scan of kScanWg values; returns the block total.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t& excl) {
    __shared__ uint32_t wsum[kScanWg / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (int q = 0; q < kScanWg / 64; q++) {
        if (q < w) base += wsum[q];
        total += wsum[q];
    }
    excl = base + inc - v;
    __syncthreads();
    return total;
}

__global__ __launch_bounds__(kScanWg) void k_scan_tiles(const uint32_t* __restrict__ in,
                                                        uint32_t* __restrict__ out, uint64_t n,
                                                        uint32_t* __restrict__ tile_sums) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t vals[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
        vals[q] = (base + q < n) ? in[base + q] : 0u;
        s += vals[q];
    }
    uint32_t excl;
    uint32_t total = block_excl_scan(s, excl);
    uint32_t run = excl;
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
        if (base + q < n) out[base + q] = run;
        run += vals[q];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Adds each tile's offset; with total, the last tile also writes the grand
// total (its offset plus its own sum).
__global__ __launch_bounds__(kScanWg) void k_scan_add(uint32_t* __restrict__ out, uint64_t n,
                                                      const uint32_t* __restrict__ tile_offs,
                                                      const uint32_t* __restrict__ tile_sums,
                                                      uint32_t* __restrict__ total) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    const uint32_t add = tile_offs[blockIdx.x];
    if (total && threadIdx.x == 0 && blockIdx.x + 1 == gridDim.x) *total = add + tile_sums[blockIdx.x];
    for (int q = threadIdx.x; q < kScanTile; q += kScanWg)
        if (base + q < n) out[base + q] += add;
}

// total = last exclusive prefix + the saved last input element (last_in holds
// exactly one value: the input may have been overwritten by an in-place scan)

constexpr int kProbeU = 8;  // gathers in flight per thread in the probe below

// Random-request rate probe: each thread gathers kProbeU float2 rows of a
// `rows`-row table at uniform pseudo-random indices (a 32-bit mix of the
// gather's number, scaled to [0, rows) by a multiply-high), all of them in
// flight together, and stores their sum so that no load is dropped.
__global__ __launch_bounds__(256) void k_probe_gather(const float2* __restrict__ table,
                                                      uint32_t rows, uint64_t threads,
                                                      uint32_t salt, float2* __restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= threads) return;
    uint32_t id[kProbeU];
#pragma unroll
    for (int u = 0; u < kProbeU; u++) {
        uint32_t h = (uint32_t)(t * kProbeU + u) * 0x9E3779B1u ^ salt * 0x85EBCA77u;
        h ^= h >> 16;
        h *= 0x7FEB352Du;
        h ^= h >> 15;
        h *= 0x846CA68Bu;
        h ^= h >> 16;
        id[u] = (uint32_t)(((uint64_t)h * rows) >> 32);
    }
    float2 v[kProbeU];
#pragma unroll
    for (int u = 0; u < kProbeU; u++) v[u] = table[id[u]];
    float2 acc = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int u = 0; u < kProbeU; u++) {
        acc.x += v[u].x;
        acc.y += v[u].y;
    }
    out[t] = acc;
}

// The same pattern as FP32 atomic adds without return (the epoch kernel's
// gradient scatter): kProbeU per thread to the x of uniform pseudo-random rows.
__global__ __launch_bounds__(256) void k_probe_atomic(float2* __restrict__ table, uint32_t rows,
                                                      uint64_t threads, uint32_t salt) {
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= threads) return;
#pragma unroll
    for (int u = 0; u < kProbeU; u++) {
        uint32_t h = (uint32_t)(t * kProbeU + u) * 0x9E3779B1u ^ salt * 0x85EBCA77u;
        h ^= h >> 16;
        h *= 0x7FEB352Du;
        h ^= h >> 15;
        h *= 0x846CA68Bu;
        h ^= h >> 16;
        unsafeAtomicAdd(&table[(uint32_t)(((uint64_t)h * rows) >> 32)].x, 1e-9f);
    }
}

// Stable LSD radix sort of u32 keys (8-bit digits) carrying u32 values.  A
// workgroup of kSortWg threads owns a tile of kSortTile consecutive keys.
constexpr int kSortWg = 256;
constexpr int kSortRounds = 16;
constexpr uint32_t kSortTile = (uint32_t)kSortWg * kSortRounds;
static_assert(kSortWg == 256, "one thread per digit value");

// counts[d * tiles + t] = number of keys of tile t whose digit is d
__global__ __launch_bounds__(kSortWg) void k_digit_counts(const uint32_t* __restrict__ key, uint32_t n,
                                                      int shift, uint32_t tiles,
                                                      uint32_t* __restrict__ counts) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; r++) {
        const uint32_t i = base + (uint32_t)r * kSortWg + threadIdx.x;
        if (i < n) atomicAdd(&h[(key[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// One stable scatter pass.  A key's place in its tile's digit-sorted order is
// the digit's start in the tile plus the number of earlier keys of the tile
// with the same digit.  Keys are taken in index order (round, then wave, then
// lane); the lanes of a wave that share a digit are found with one ballot per
// digit bit.  The tile is first sorted into LDS, then written out run by run,
// so consecutive threads store consecutive places of a digit's run (a direct
// scatter stored each lane's key to its own place, up to 64 places per
// wavefront store).  A digit's runs lie at offs (exclusive scan of counts,
// digit-major).  val_in == nullptr stands for the identity (i).
__global__ __launch_bounds__(kSortWg) void k_digit_scatter(const uint32_t* __restrict__ key_in,
                                                       const uint32_t* __restrict__ val_in,
                                                       uint32_t n, int shift, uint32_t tiles,
                                                       const uint32_t* __restrict__ offs,
                                                       uint32_t* __restrict__ key_out,
                                                       uint32_t* __restrict__ val_out) {
    constexpr int kWaves = kSortWg / 64;
    __shared__ uint32_t run[256];           // keys of each digit placed so far in this tile
    __shared__ uint32_t wpre[kWaves][256];  // per round: wave counts, then wave offsets
    __shared__ uint32_t lstart[256];        // start of each digit's run in the tile
    __shared__ uint32_t gstart[256];        // start of that run in the output
    __shared__ uint32_t sk[kSortTile], sv[kSortTile];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t base = blockIdx.x * kSortTile;
    const uint32_t len = n - base < kSortTile ? n - base : kSortTile;
    {  // this tile's count of digit tid, from consecutive scanned offsets
        const size_t idx = (size_t)tid * tiles + blockIdx.x;
        const uint32_t o = offs[idx];
        const uint32_t nx = idx + 1 < (size_t)256 * tiles ? offs[idx + 1] : n;
        gstart[tid] = o;
        lstart[tid] = nx - o;
        run[tid] = 0u;
    }
    __syncthreads();
    const uint32_t own = lstart[tid];
    for (int off = 1; off < 256; off <<= 1) {  // inclusive scan of the tile's counts
        const uint32_t t = tid >= off ? lstart[tid - off] : 0u;
        __syncthreads();
        lstart[tid] += t;
        __syncthreads();
    }
    lstart[tid] -= own;
    for (int r = 0; r < kSortRounds; r++) {
#pragma unroll
        for (int w = 0; w < kWaves; w++) wpre[w][tid] = 0u;
        __syncthreads();
        const uint32_t i = base + (uint32_t)r * kSortWg + tid;
        const bool valid = i < n;
        const uint32_t k = valid ? key_in[i] : 0u;
        const uint32_t d = (k >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        if (valid && rank == 0) wpre[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {  // thread tid owns digit tid: wave offsets on top of the tile's running count
            uint32_t acc = run[tid];
#pragma unroll
            for (int w = 0; w < kWaves; w++) {
                const uint32_t c = wpre[w][tid];
                wpre[w][tid] = acc;
                acc += c;
            }
            run[tid] = acc;
        }
        __syncthreads();
        if (valid) {
            const uint32_t lp = lstart[d] + wpre[wave][d] + rank;
            sk[lp] = k;
            sv[lp] = val_in ? val_in[i] : i;
        }
        __syncthreads();
    }
    for (uint32_t q = tid; q < len; q += kSortWg) {
        const uint32_t k = sk[q];
        const uint32_t d = (k >> shift) & 255u;
        const uint32_t pos = gstart[d] + (q - lstart[d]);
        key_out[pos] = k;
        val_out[pos] = sv[q];
    }
}

}  // namespace

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total_dev,
                        hipStream_t s) {
    if (n == 0) {
        if (total_dev) MRS_HIP(hipMemsetAsync(total_dev, 0, sizeof(uint32_t), s));
        return;
    }
    const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 1) {
        // one tile: its sum is the total
        DevBuf<uint32_t> sum1(total_dev ? 0 : 1);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanWg), 0, s, in, out, n,
                           total_dev ? total_dev : sum1.p);
        MRS_CHECK_LAUNCH();
        if (sum1.freed_at_once()) MRS_HIP(hipStreamSynchronize(s));
        return;
    }
    DevBuf<uint32_t> sums(tiles), offs(tiles);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(kScanWg), 0, s, in, out, n,
                       sums.p);
    MRS_CHECK_LAUNCH();
    exclusive_scan_u32(sums.p, offs.p, tiles, nullptr, s);
    // the last tile writes the total: its offset plus its sum (sums is unchanged)
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(kScanWg), 0, s, out, n, offs.p,
                       sums.p, total_dev);
    MRS_CHECK_LAUNCH();
    // temporaries are freed on return: unless that free is stream-ordered it
    // is immediate, so the queued work must be finished first
    if (sums.freed_at_once()) MRS_HIP(hipStreamSynchronize(s));
}

void radix_sort_pairs_u32(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t n,
                          int key_bits, uint32_t* keys_out, uint32_t* vals_out, hipStream_t s) {
    if (n == 0) return;
    const int passes = std::max(1, (key_bits + 7) / 8);
    const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
    DevBuf<uint32_t> counts((size_t)256 * tiles), kt(n), vt(n);
    // the last pass writes (keys_out, vals_out); earlier ones alternate with
    // the temporaries so that no pass reads the buffer it writes
    const uint32_t* kin = keys_in;
    const uint32_t* vin = vals_in;
    for (int pass = 0; pass < passes; pass++) {
        const bool last = ((passes - 1 - pass) & 1) == 0;
        uint32_t* kout = last ? keys_out : kt.p;
        uint32_t* vout = last ? vals_out : vt.p;
        hipLaunchKernelGGL(k_digit_counts, dim3(tiles), dim3(kSortWg), 0, s, kin, n, 8 * pass,
                           tiles, counts.p);
        MRS_CHECK_LAUNCH();
        exclusive_scan_u32(counts.p, counts.p, (uint64_t)256 * tiles, nullptr, s);
        hipLaunchKernelGGL(k_digit_scatter, dim3(tiles), dim3(kSortWg), 0, s, kin, vin, n,
                           8 * pass, tiles, counts.p, kout, vout);
        MRS_CHECK_LAUNCH();
        kin = kout;
        vin = vout;
    }
    // temporaries are freed on return
    if (counts.freed_at_once()) MRS_HIP(hipStreamSynchronize(s));
}

void ensure_dynamic_lds(const void* kern, size_t bytes) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> limit;
    int dev = 0;
    MRS_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    size_t& cur = limit[{dev, kern}];
    if (bytes <= cur) return;
    MRS_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    cur = bytes;
}

}  // namespace mrs

// ------------------------------------------------------------------ C ABI

extern "C" {

const char* mrs_last_error(void) { return mrs::g_last_error.c_str(); }

int mrs_ctx_create(int32_t device, size_t hbm_budget, mrs_ctx** out) {
    return mrs::guarded([&] {
        if (!out) throw mrs::Error(MRS_INVALID_ARGUMENT, "out is NULL");
        int count = 0;
        MRS_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count)
            throw mrs::Error(MRS_DEVICE_ERROR, "no such HIP device: " + std::to_string(device));
        MRS_HIP(hipSetDevice(device));
        // keep freed pool memory for reuse instead of returning it at each sync
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
        auto* c = new mrs_ctx();
        c->device = device;
        c->hbm_budget = hbm_budget;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw mrs::Error(MRS_DEVICE_ERROR, "hipStreamCreate failed");
        }
        c->pool = std::make_shared<mrs::DevicePool>(device, c->stream);
        *out = c;
    });
}

int mrs_ctx_destroy(mrs_ctx* ctx) {
    return mrs::guarded([&] {
        if (!ctx) return;
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int mrs_ctx_synchronize(mrs_ctx* ctx) {
    return mrs::guarded([&] {
        if (!ctx) throw mrs::Error(MRS_INVALID_ARGUMENT, "ctx is NULL");
        MRS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

void* mrs_ctx_stream(mrs_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

size_t mrs_ctx_cached_bytes(mrs_ctx* ctx) {
    return ctx && ctx->pool ? ctx->pool->cached_bytes() : 0;
}

int mrs_ctx_release_cache(mrs_ctx* ctx) {
    return mrs::guarded([&] {
        if (!ctx) throw mrs::Error(MRS_INVALID_ARGUMENT, "ctx is NULL");
        MRS_HIP(hipSetDevice(ctx->device));
        if (ctx->pool) ctx->pool->trim();
    });
}

namespace {

// ops per second of `reps` launches of the gather (atomic = false) or the
// atomic probe, after two warm-up launches
int probe_random(mrs_ctx* ctx, int64_t rows, int64_t gathers, int32_t reps, double* per_s,
                 bool atomic) {
    return mrs::guarded([&] {
        if (!ctx || !per_s) throw mrs::Error(MRS_INVALID_ARGUMENT, "NULL argument");
        if (rows < 1 || rows > (int64_t)UINT32_MAX)
            throw mrs::Error(MRS_INVALID_ARGUMENT, "rows must be in [1, 2^32)");
        if (gathers < mrs::kProbeU || gathers > ((int64_t)1 << 36) || reps < 1)
            throw mrs::Error(MRS_INVALID_ARGUMENT, "gathers must be in [8, 2^36], reps >= 1");
        MRS_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        mrs::StreamScope scope(s, ctx->pool);
        const uint64_t threads = (uint64_t)gathers / mrs::kProbeU;
        mrs::DevBuf<float2> table((size_t)rows), out((size_t)threads);
        MRS_HIP(hipMemsetAsync(table.p, 0, table.bytes(), s));
        const dim3 grid((unsigned)((threads + 255) / 256));
        auto launch = [&](uint32_t salt) {
            if (atomic)
                hipLaunchKernelGGL(mrs::k_probe_atomic, grid, dim3(256), 0, s, table.p,
                                   (uint32_t)rows, threads, salt);
            else
                hipLaunchKernelGGL(mrs::k_probe_gather, grid, dim3(256), 0, s, table.p,
                                   (uint32_t)rows, threads, salt, out.p);
        };
        for (int w = 0; w < 2; w++) launch((uint32_t)w);
        MRS_CHECK_LAUNCH();
        hipEvent_t e0 = nullptr, e1 = nullptr;
        MRS_HIP(hipEventCreate(&e0));
        MRS_HIP(hipEventCreate(&e1));
        float ms = 0.0f;
        hipError_t err = hipEventRecord(e0, s);
        for (int r = 0; r < reps && err == hipSuccess; r++) launch((uint32_t)(r + 2));
        if (err == hipSuccess) err = hipGetLastError();
        if (err == hipSuccess) err = hipEventRecord(e1, s);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        MRS_HIP(err);
        *per_s = ms > 0.0f ? (double)threads * mrs::kProbeU * reps / (ms * 1e-3) : 0.0;
    });
}

}  // namespace

int mrs_probe_random_gather(mrs_ctx* ctx, int64_t rows, int64_t gathers, int32_t reps,
                            double* per_s) {
    return probe_random(ctx, rows, gathers, reps, per_s, false);
}

int mrs_probe_random_atomic(mrs_ctx* ctx, int64_t rows, int64_t ops, int32_t reps,
                            double* per_s) {
    return probe_random(ctx, rows, ops, reps, per_s, true);
}

int mrs_ctx_set_knn_pruning(mrs_ctx* ctx, int32_t on) {
    return mrs::guarded([&] {
        if (!ctx) throw mrs::Error(MRS_INVALID_ARGUMENT, "ctx is NULL");
        ctx->knn_prune = on > 0 ? 1 : (on < 0 ? -1 : 0);
    });
}

int mrs_ctx_set_knn_work_chunk(mrs_ctx* ctx, int32_t stages) {
    return mrs::guarded([&] {
        if (!ctx) throw mrs::Error(MRS_INVALID_ARGUMENT, "ctx is NULL");
        if (stages < 0) throw mrs::Error(MRS_INVALID_ARGUMENT, "stages must be >= 0");
        ctx->knn_chunk = (uint32_t)stages;
    });
}

static std::string lower(const char* s) {
    std::string r = s ? s : "";
    std::transform(r.begin(), r.end(), r.begin(), [](unsigned char c) { return std::tolower(c); });
    return r;
}

int32_t mrs_parse_umap_optimiser_gpu(const char* s) {
    std::string t = lower(s);
    if (t == "adam_gpu") return 0;
    if (t == "adam_parallel") return 1;
    if (t == "adam") return 2;
    if (t == "sgd") return 3;
    return -1;
}

uint32_t mrs_hash_neg(uint32_t seed, uint32_t node, uint32_t epoch, uint32_t edge_local,
                      uint32_t neg, uint32_t n) {
    if (n == 0) return 0;
    return mrs::hash_neg(seed, node, epoch, edge_local, neg, mrs::make_fastmod(n));
}

int32_t mrs_parse_ann_search_gpu(const char* s) {
    std::string t = lower(s);
    if (t == "ivf_gpu" || t == "ivf") return 0;
    if (t == "nndescent_gpu" || t == "nndescent") return 1;
    if (t == "exhaustive_gpu" || t == "exhaustive") return 2;
    return -1;
}

}  // extern "C"
