// Fit server: continuous batching of many uneven UMAP fits on one GPU.
//
// The reference has no serving layer; this is the request loop its "adam_gpu"
// fits are batched under.  A request is one umap_gpu fit (src/lib.rs:3465-3487
// semantics, same parameters and seed behaviour).  The server:
//
//   * admits requests in arrival order while their estimated device memory
//     fits the HBM budget (and a cap on resident fits);
//   * prepares admitted fits on P worker threads, each with its own stream
//     (kNN -> fuzzy graph -> spectral layout -> optimiser state), so the
//     preparations of several fits overlap on the device;
//   * advances every resident fit by one epoch per tick on one batcher
//     stream: one segmented edge kernel and one segmented Adam kernel cover
//     all resident edge-parallel fits, each at its own epoch (a block finds its
//     fit in a table of fit descriptors uploaded per tick); deterministic
//     (node-exact) fits run their own per-fit launches in the same tick;
//   * retires a fit after its last epoch (transpose to [n_dim][n]; the caller
//     collects it with mrs_server_wait) and frees its memory in stream order.
//
// Memory: a prep worker builds a fit's optimiser state from its own context's
// block cache, and the batcher uses and releases it on the batcher stream.
// The cache fences such cross-stream reuse (block_cache.hpp): a block the
// batcher returns carries an event recorded after the fit's last queued
// kernel, and the prep stream that takes it next waits for that event.
//
// A fit's epochs depend only on its own state, seed and epoch counter (the
// negative-sample hash is keyed by node, epoch and CSR slot), so batching
// changes nothing for the node-exact strategy and only the FP32 atomic
// summation order for the edge-parallel one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "epoch_device.hpp"
#include "fit.hpp"

namespace mrs {

namespace {

constexpr int kWg = 256;

}  // namespace

namespace {

// One resident edge-parallel fit in a batched epoch launch.
struct FitDesc {
    const uint4* edges;
    const float* per;
    float* next;
    float* embd;
    float* grad;
    float* m;
    float* v;
    uint32_t n_edges, n;
    uint32_t epoch;
    float epoch_f;
    FastMod fm;
    uint32_t seed;
    int32_t neg_rate;
    float a, b, two_a_b, two_gamma_b;
    float fused_lr, epsc, omb1, omb2;
    uint32_t blk_node0;  // first block of this fit in the node launch
};

// prep worker threads per server (the most that have run without fault)
constexpr int kMaxPrepWorkers = 8;

// Workgroups of a launch go to the chip's 8 XCDs round-robin: block b runs on
// XCD b mod 8.  The edge launch gives each XCD its own bin of edge blocks, so
// that a fit's edges run on as few XCDs as possible and its embedding, the
// target of every negative-sample gather, stays in those XCDs' L2 (each XCD
// has its own).  Bins are filled in fit order to equal size; a fit larger
// than the space left in a bin continues in the next.  A piece is one fit's
// run of blocks inside one bin.
constexpr uint32_t kXcds = 8;
struct Piece {
    uint32_t fit;    // index into the launch's descriptors
    uint32_t blk0;   // the piece's first edge block within its fit
    uint32_t slot0;  // its first slot in the bin
    uint32_t count;  // blocks
};
struct Bins {
    uint32_t first[kXcds + 1];  // pieces of bin x: [first[x], first[x + 1]), by slot
};

// index of the fit owning block `blk` of the node launch (fits sorted by first
// block); uniform per block
__device__ inline int fit_of_block(const FitDesc* __restrict__ fits, int n_fits, uint32_t blk) {
    int lo = 0, hi = n_fits - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (fits[mid].blk_node0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One epoch of every fit of the launch: block b takes slot b / 8 of bin b % 8,
// i.e. kSegBlockEdges consecutive edges of one fit, as four wavefronts of
// kSegSpan-edge spans with the firing edges compacted (epoch_edge_span, the
// single-fit kernel's code).  About one edge in five fires per epoch, so a
// 64-edge span left a wavefront ~13 firing edges for 64 lanes and the kernel
// VALU-bound (75% VALU busy on the served stream); 256-edge spans fill the
// lanes.  Slots past the end of their bin exit at once (the whole block).
constexpr int kSegSpan = 256;
constexpr uint32_t kSegBlockEdges = (kWg / 64) * kSegSpan;
template <int DIM>
__global__ __launch_bounds__(kWg) void k_seg_epoch_edges(const FitDesc* __restrict__ fits,
                                                         const Piece* __restrict__ pieces,
                                                         Bins bins) {
    constexpr int kWaves = kWg / 64;
    __shared__ SpanQueue<kSegSpan> queue[kWaves];
    const uint32_t x = blockIdx.x % kXcds, t = blockIdx.x / kXcds;
    uint32_t lo = bins.first[x], hi = bins.first[x + 1];
    if (lo == hi) return;
    while (hi - lo > 1) {  // the last piece of the bin starting at or before slot t
        const uint32_t mid = (lo + hi) >> 1;
        if (pieces[mid].slot0 <= t) lo = mid;
        else hi = mid;
    }
    const Piece pc = pieces[lo];
    if (t < pc.slot0 || t - pc.slot0 >= pc.count) return;
    const FitDesc& f = fits[pc.fit];
    const int wave = threadIdx.x >> 6;
    const uint64_t base = ((uint64_t)(pc.blk0 + (t - pc.slot0)) * kWaves + wave) * kSegSpan;
    epoch_edge_span<DIM, false, kSegSpan>(base, queue[wave], f.edges, f.per, f.next, f.embd, f.grad,
                                    f.n_edges, f.epoch, f.epoch_f, f.seed, f.fm, f.a, f.b,
                                    f.two_a_b, f.two_gamma_b, f.neg_rate, nullptr, nullptr,
                                    nullptr);
}

template <int DIM>
__global__ __launch_bounds__(kWg) void k_seg_adam_clear(const FitDesc* __restrict__ fits,
                                                        int n_fits) {
    const FitDesc& f = fits[fit_of_block(fits, n_fits, blockIdx.x)];
    const uint32_t node = (blockIdx.x - f.blk_node0) * kWg + threadIdx.x;
    if (node >= f.n) return;
    adam_clear_node<DIM>(node, f.grad, f.m, f.v, f.embd, f.fused_lr, f.epsc, f.omb1, f.omb2,
                         nullptr);
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

bool edge_dim_ok(int dim) { return dim == 1 || dim == 2 || dim == 3 || dim == 4 || dim == 8; }

}  // namespace

// Peak device bytes of one fit's preparation (kNN operands, candidate lists and
// queues, graph construction, Lanczos basis) and its resident state, per point
// plus fixed overhead: an upper estimate used for admission.
size_t fit_bytes_estimate(int64_t n, int32_t d, int32_t k, int32_t n_dim) {
    const size_t np = (size_t)((n + 255) / 256) * 256;
    const size_t knn = np * (size_t)d * 6 + np * 64 * 8 * 2 + (size_t)n * k * 8;
    const size_t graph = (size_t)n * k * 64;
    const size_t lanczos = (size_t)n * (4 * ((size_t)n_dim + 1) + 30 + 8) * 8 + (size_t)n * k * 16;
    const size_t state = (size_t)n * k * 40 + (size_t)n * n_dim * 20;
    return (size_t)n * d * 4 + std::max(knn, std::max(graph, lanczos)) + state + (64u << 20);
}

}  // namespace mrs

enum ReqState { kQueued = 0, kPreparing = 1, kEpochs = 2, kDone = 3, kFailed = 4 };

struct mrs_server_request {
    int64_t id = 0;
    const float* data = nullptr;  // caller memory, valid until the request completes
    int64_t n = 0;
    int32_t d = 0;
    mrs_umap_params_gpu params{};
    std::string optimiser, ann_type, initialisation, dist_metric;
    uint64_t seed = 0;
    size_t est_bytes = 0;
    std::atomic<int> state{kQueued};
    int status = MRS_OK;
    std::string error;
    double t_submit = 0, t_admit = 0, t_ready = 0, t_done = 0;
    std::unique_ptr<mrs_umap_state> st;
    int32_t epoch = 0;
    mrs::DevBuf<float> out;  // [n_dim][n] after the last epoch
    hipEvent_t done_ev = nullptr;
};

struct mrs_server {
    int device = 0;
    size_t budget = 0;
    int max_active = 0;
    std::vector<mrs_ctx*> prep_ctx;
    mrs_ctx* batch_ctx = nullptr;
    std::vector<std::thread> workers;
    std::thread batcher;

    std::mutex mu;
    std::condition_variable cv_work, cv_batch, cv_done;
    bool stopping = false;
    int64_t next_id = 1;
    size_t in_use = 0, peak_in_use = 0;
    int resident = 0, peak_resident = 0;
    std::deque<mrs_server_request*> pending;  // submitted, not yet admitted
    std::deque<mrs_server_request*> ready;    // prepared, waiting for the batcher
    std::map<int64_t, std::unique_ptr<mrs_server_request>> all;
    int64_t ticks = 0, completed = 0, failed = 0;

    // per-tick descriptor upload: a ring of pinned host tables and of device
    // tables (max_active entries per slot), one event per tick in flight; a
    // tick never writes the table the previous tick's kernels read
    // A slot holds the tick's descriptors followed by its pieces (at most
    // max_active + 8 per dimension group), in units of FitDesc.
    static constexpr int kRing = 8;
    mrs::FitDesc* host_desc[kRing] = {nullptr};
    hipEvent_t tick_ev[kRing] = {nullptr};
    mrs::DevBuf<mrs::FitDesc> dev_desc;
    size_t slot_units = 0;
};

namespace mrs {
namespace {

void finish(mrs_server* sv, mrs_server_request* r, int status, const std::string& err) {
    std::lock_guard<std::mutex> lk(sv->mu);
    r->status = status;
    r->error = err;
    r->t_done = now_ms();
    r->state = status == MRS_OK ? kDone : kFailed;
    if (status == MRS_OK) sv->completed++;
    else sv->failed++;
    sv->in_use -= r->est_bytes;
    sv->resident--;
    sv->cv_done.notify_all();
    sv->cv_work.notify_all();   // budget freed: admission may proceed
    sv->cv_batch.notify_all();  // a drained, stopping server lets the batcher exit
}

// Prep worker: admit the oldest pending request once it fits the budget,
// prepare it on this worker's stream and hand it to the batcher.
void worker_loop(mrs_server* sv, int w) {
    mrs_ctx* ctx = sv->prep_ctx[w];
    (void)hipSetDevice(sv->device);
    for (;;) {
        mrs_server_request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_work.wait(lk, [&] {
                if (sv->stopping && sv->pending.empty()) return true;
                if (sv->pending.empty()) return false;
                mrs_server_request* h = sv->pending.front();
                const bool fits_budget = sv->resident == 0 || sv->in_use + h->est_bytes <= sv->budget;
                return fits_budget && sv->resident < sv->max_active;
            });
            if (sv->pending.empty()) return;  // stopping
            r = sv->pending.front();
            sv->pending.pop_front();
            sv->in_use += r->est_bytes;
            sv->peak_in_use = std::max(sv->peak_in_use, sv->in_use);
            sv->resident++;
            sv->peak_resident = std::max(sv->peak_resident, sv->resident);
            r->t_admit = now_ms();
            r->state = kPreparing;
        }
        int status = MRS_OK;
        std::string err;
        try {
            // every buffer of the preparation comes from this worker's cache
            StreamScope scope(ctx->stream, ctx->pool);
            FitTimings t;
            r->st = fit_prepare(ctx, r->data, r->n, r->d, nullptr, nullptr, 0, &r->params,
                                nullptr, r->seed, 0, t);
            MRS_HIP(hipStreamSynchronize(ctx->stream));
            // keep the cache to a share of the budget (blocks of many sizes pile up)
            if (ctx->pool->cached_bytes() > sv->budget / (4 * sv->prep_ctx.size()))
                ctx->pool->trim();
        } catch (const Error& e) {
            status = e.status;
            err = e.what();
        } catch (const std::exception& e) {
            status = MRS_DEVICE_ERROR;
            err = e.what();
        }
        if (status != MRS_OK) {
            if (r->st) {
                StreamScope scope(ctx->stream, ctx->pool);
                r->st.reset();
            }
            finish(sv, r, status, err);
            continue;
        }
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            r->t_ready = now_ms();
            r->state = kEpochs;
            sv->ready.push_back(r);
        }
        sv->cv_batch.notify_one();
    }
}

void fill_desc(FitDesc& f, const mrs_umap_state* st, int32_t epoch, uint32_t seed) {
    const mrs_umap_optim_params& p = st->params;
    const AdamScalars sc = adam_scalars(p, epoch);
    f.edges = st->edges.p;
    f.per = st->per.p;
    f.next = st->next.p;
    f.embd = st->embd.p;
    f.grad = st->grad.p;
    f.m = st->m.p;
    f.v = st->v.p;
    f.n_edges = (uint32_t)st->n_edges;
    f.n = (uint32_t)st->n;
    f.epoch = (uint32_t)epoch;
    f.epoch_f = (float)epoch;
    f.fm = make_fastmod((uint32_t)st->n);
    f.seed = seed;
    f.neg_rate = p.neg_sample_rate;
    f.a = p.a;
    f.b = p.b;
    f.two_a_b = 2.0f * p.a * p.b;
    f.two_gamma_b = 2.0f * p.gamma * p.b;
    f.fused_lr = sc.fused_lr;
    f.epsc = sc.epsc;
    f.omb1 = 1.0f - p.beta1;
    f.omb2 = 1.0f - p.beta2;
}

// Batcher: one epoch of every resident fit per tick.
void batcher_loop(mrs_server* sv) {
    mrs_ctx* ctx = sv->batch_ctx;
    hipStream_t s = ctx->stream;
    (void)hipSetDevice(sv->device);
    // outputs come from the batcher's cache; states released here go back to
    // their prep worker's cache fenced by an event on this stream
    StreamScope scope(s, ctx->pool);
    std::vector<mrs_server_request*> active;
    struct Group {
        int dim, first, count;
        uint32_t blocks_e, blocks_n;
        uint32_t piece0, slots;  // the group's first piece; slots per XCD bin
        Bins bins;
    };
    int64_t tick = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_batch.wait(lk, [&] {
                return !sv->ready.empty() || !active.empty() ||
                       (sv->stopping && sv->pending.empty() && sv->resident == 0);
            });
            while (!sv->ready.empty()) {
                mrs_server_request* r = sv->ready.front();
                sv->ready.pop_front();
                r->st->ctx = ctx;  // the state now lives on the batcher stream
                active.push_back(r);
            }
            if (active.empty()) return;  // stopping and drained
        }
        try {
            // at most kRing ticks in flight: the slot's previous tick (and its
            // descriptor copy out of this host buffer) has completed
            const int slot = (int)(tick % mrs_server::kRing);
            MRS_HIP(hipEventSynchronize(sv->tick_ev[slot]));
            FitDesc* hd = sv->host_desc[slot];
            // edge-parallel fits, grouped by dimension: one segmented launch pair per group
            std::vector<Group> groups;
            int nd = 0;
            for (int dim : {1, 2, 3, 4, 8}) {
                Group g{dim, nd, 0, 0, 0, 0, 0, {}};
                for (mrs_server_request* r : active) {
                    mrs_umap_state* st = r->st.get();
                    if (r->params.epoch_strategy != MRS_EPOCH_EDGE_ATOMIC || st->dim != dim)
                        continue;
                    // a new fit's gradients to the -0.0 markers, on this stream
                    if (!st->grad_clean) clear_update_marks(st);
                    FitDesc& f = hd[nd++];
                    fill_desc(f, st, r->epoch, (uint32_t)r->seed);
                    f.blk_node0 = g.blocks_n;
                    g.blocks_e += div_up(st->n_edges, kSegBlockEdges);
                    g.blocks_n += div_up(st->n, kWg);
                }
                g.count = nd - g.first;
                if (g.count) groups.push_back(g);
            }
            if (nd) {
                // the pieces follow the descriptors in the same table
                Piece* hp = reinterpret_cast<Piece*>(hd + nd);
                uint32_t np = 0;
                for (Group& g : groups) {
                    g.piece0 = np;
                    g.slots = div_up(g.blocks_e, kXcds);
                    uint32_t x = 0, filled = 0, cnt = 0;
                    g.bins.first[0] = 0;
                    for (int q = 0; q < g.count; q++) {
                        uint32_t left = div_up(hd[g.first + q].n_edges, kSegBlockEdges), blk0 = 0;
                        while (left) {
                            const uint32_t take = std::min(left, g.slots - filled);
                            hp[np + cnt++] = Piece{(uint32_t)q, blk0, filled, take};
                            blk0 += take;
                            left -= take;
                            filled += take;
                            if (filled == g.slots) {
                                g.bins.first[++x] = cnt;
                                filled = 0;
                            }
                        }
                    }
                    for (uint32_t q = x + 1; q <= kXcds; q++) g.bins.first[q] = cnt;
                    np += cnt;
                }
                const size_t bytes = sizeof(FitDesc) * nd + sizeof(Piece) * np;
                if (bytes > sizeof(FitDesc) * sv->slot_units)
                    throw Error(MRS_DEVICE_ERROR, "fit server: tick table overflow");
                FitDesc* dtab = sv->dev_desc.p + (size_t)slot * sv->slot_units;
                MRS_HIP(hipMemcpyAsync(dtab, hd, bytes, hipMemcpyHostToDevice, s));
                const Piece* dpieces = reinterpret_cast<const Piece*>(dtab + nd);
                for (const Group& g : groups) {
                    const FitDesc* dd = dtab + g.first;
                    auto go = [&](auto ke, auto ka) {
                        hipLaunchKernelGGL(ke, dim3(g.slots * kXcds), dim3(kWg), 0, s, dd,
                                           dpieces + g.piece0, g.bins);
                        hipLaunchKernelGGL(ka, dim3(g.blocks_n), dim3(kWg), 0, s, dd, g.count);
                    };
                    switch (g.dim) {
                        case 1: go(k_seg_epoch_edges<1>, k_seg_adam_clear<1>); break;
                        case 2: go(k_seg_epoch_edges<2>, k_seg_adam_clear<2>); break;
                        case 3: go(k_seg_epoch_edges<3>, k_seg_adam_clear<3>); break;
                        case 4: go(k_seg_epoch_edges<4>, k_seg_adam_clear<4>); break;
                        default: go(k_seg_epoch_edges<8>, k_seg_adam_clear<8>); break;
                    }
                    MRS_CHECK_LAUNCH();
                }
            }
            // node-exact (and any other) fits: their own launches in the same tick
            for (mrs_server_request* r : active) {
                mrs_umap_state* st = r->st.get();
                const bool batched = r->params.epoch_strategy == MRS_EPOCH_EDGE_ATOMIC &&
                                     edge_dim_ok(st->dim);
                if (!batched)
                    run_epochs(st, r->epoch, 1, (uint32_t)r->seed, r->params.epoch_strategy);
            }
            MRS_HIP(hipEventRecord(sv->tick_ev[slot], s));
            for (mrs_server_request* r : active) r->epoch++;
            // retire finished fits: [n_dim][n] output on the device, state freed
            // in stream order.  A finished request leaves `active` before it is
            // handed back (a waiter may free it at once), and each is finished
            // exactly once, with its own error if its retirement fails.
            std::vector<mrs_server_request*> keep, done;
            for (mrs_server_request* r : active)
                (r->epoch < r->st->params.n_epochs ? keep : done).push_back(r);
            active.swap(keep);
            for (mrs_server_request* r : done) {
                int status = MRS_OK;
                std::string err;
                try {
                    r->out.alloc((size_t)r->st->n * r->st->dim);
                    fit_output_async(r->st.get(), r->out.p);
                    MRS_HIP(hipEventCreateWithFlags(&r->done_ev, hipEventDisableTiming));
                    MRS_HIP(hipEventRecord(r->done_ev, s));
                } catch (const std::exception& e) {
                    const Error* me = dynamic_cast<const Error*>(&e);
                    status = me ? me->status : MRS_DEVICE_ERROR;
                    err = e.what();
                }
                r->st.reset();
                finish(sv, r, status, err);
            }
            tick++;
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->ticks = tick;
        } catch (const std::exception& e) {
            const Error* me = dynamic_cast<const Error*>(&e);
            const int status = me ? me->status : MRS_DEVICE_ERROR;
            for (mrs_server_request* r : active) {
                r->st.reset();
                finish(sv, r, status, e.what());
            }
            active.clear();
        }
    }
}

}  // namespace
}  // namespace mrs

// ------------------------------------------------------------------ C ABI

using mrs::Error;

extern "C" {

size_t mrs_fit_bytes_estimate(int64_t n, int32_t d, int32_t k, int32_t n_dim) {
    if (n <= 0 || d <= 0 || k <= 0 || n_dim <= 0) return 0;
    return mrs::fit_bytes_estimate(n, d, k, n_dim);
}

int mrs_server_destroy(mrs_server* sv);

int mrs_server_create(int32_t device, size_t hbm_budget, int32_t prep_workers, int32_t max_active,
                      mrs_server** out) {
    return mrs::guarded([&] {
        if (!out) throw Error(MRS_INVALID_ARGUMENT, "out is NULL");
        // Up to 8 workers have run cleanly; one run with 12 workers (and 32
        // hardware queues) ended in a device memory fault whose cause has not
        // been found, so larger counts are refused until it is.
        if (prep_workers < 1 || prep_workers > mrs::kMaxPrepWorkers)
            throw Error(MRS_INVALID_ARGUMENT,
                        "prep_workers must be 1.." + std::to_string(mrs::kMaxPrepWorkers));
        if (max_active < 1 || max_active > 4096)
            throw Error(MRS_INVALID_ARGUMENT, "max_active must be 1..4096");
        auto* sv = new mrs_server();
        sv->device = device;
        sv->max_active = max_active;
        try {
            mrs_ctx* c = nullptr;
            if (mrs_ctx_create(device, hbm_budget, &c) != MRS_OK)
                throw Error(MRS_DEVICE_ERROR, mrs_last_error());
            sv->batch_ctx = c;
            for (int w = 0; w < prep_workers; w++) {
                if (mrs_ctx_create(device, hbm_budget, &c) != MRS_OK)
                    throw Error(MRS_DEVICE_ERROR, mrs_last_error());
                sv->prep_ctx.push_back(c);
            }
            if (hbm_budget == 0) {
                size_t free_b = 0, total_b = 0;
                MRS_HIP(hipMemGetInfo(&free_b, &total_b));
                hbm_budget = free_b / 2;
            }
            sv->budget = hbm_budget;
            // descriptors, then pieces: at most one per fit plus one per bin
            // boundary in each of the five dimension groups
            sv->slot_units = (size_t)max_active +
                             mrs::div_up(sizeof(mrs::Piece) * ((size_t)max_active + 5 * mrs::kXcds),
                                         sizeof(mrs::FitDesc));
            for (int q = 0; q < mrs_server::kRing; q++) {
                MRS_HIP(hipHostMalloc((void**)&sv->host_desc[q],
                                      sizeof(mrs::FitDesc) * sv->slot_units, hipHostMallocDefault));
                MRS_HIP(hipEventCreateWithFlags(&sv->tick_ev[q], hipEventDisableTiming));
                MRS_HIP(hipEventRecord(sv->tick_ev[q], sv->batch_ctx->stream));
            }
            {
                mrs::StreamScope scope(sv->batch_ctx->stream);
                sv->dev_desc.alloc(sv->slot_units * mrs_server::kRing);
            }
            for (int w = 0; w < prep_workers; w++) sv->workers.emplace_back(mrs::worker_loop, sv, w);
            sv->batcher = std::thread(mrs::batcher_loop, sv);
        } catch (...) {
            mrs_server_destroy(sv);
            throw;
        }
        *out = sv;
    });
}

int mrs_server_submit(mrs_server* sv, const float* data, int64_t n, int32_t d,
                      const mrs_umap_params_gpu* params, uint64_t seed, int64_t* ticket) {
    return mrs::guarded([&] {
        if (!sv || !params || !ticket) throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        if (n <= 0 || d <= 0 || !data) throw Error(MRS_NO_DATA, "UMAP: empty data (NoData)");
        if (params->n_dim <= 0 || params->n_dim > 16)
            throw Error(MRS_INVALID_ARGUMENT, "n_dim must be 1..16");
        if (params->k <= 0 || params->k > n - 1) throw Error(MRS_INVALID_ARGUMENT, "k must be in [1, n-1]");
        if (params->optim_params.n_epochs <= 0)
            throw Error(MRS_INVALID_ARGUMENT, "n_epochs must be positive");
        auto r = std::make_unique<mrs_server_request>();
        r->data = data;
        r->n = n;
        r->d = d;
        r->params = *params;
        auto keep = [](std::string& dst, const char*& p) {
            dst = p ? p : "";
            p = p ? dst.c_str() : nullptr;
        };
        keep(r->optimiser, r->params.optimiser);
        keep(r->ann_type, r->params.ann_type);
        keep(r->initialisation, r->params.initialisation);
        keep(r->dist_metric, r->params.dist_metric);
        r->seed = seed;
        r->est_bytes = mrs::fit_bytes_estimate(n, d, params->k, params->n_dim);
        if (r->est_bytes > sv->budget)
            throw Error(MRS_INVALID_ARGUMENT, "request larger than the server's HBM budget");
        r->t_submit = mrs::now_ms();
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->stopping) throw Error(MRS_INVALID_ARGUMENT, "server is shutting down");
        r->id = sv->next_id++;
        *ticket = r->id;
        sv->pending.push_back(r.get());
        sv->all[r->id] = std::move(r);
        sv->cv_work.notify_all();
    });
}

int mrs_server_poll(mrs_server* sv, int64_t ticket, int32_t* state) {
    return mrs::guarded([&] {
        if (!sv || !state) throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->all.find(ticket);
        if (it == sv->all.end()) throw Error(MRS_INVALID_ARGUMENT, "unknown ticket");
        *state = it->second->state;
    });
}

int mrs_server_wait(mrs_server* sv, int64_t ticket, float* out, double* times_ms) {
    std::unique_ptr<mrs_server_request> r;
    int rc = mrs::guarded([&] {
        if (!sv) throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->all.find(ticket);
        if (it == sv->all.end()) throw Error(MRS_INVALID_ARGUMENT, "unknown ticket");
        mrs_server_request* q = it->second.get();
        sv->cv_done.wait(lk, [&] { return q->state == kDone || q->state == kFailed; });
        r = std::move(it->second);
        sv->all.erase(it);
    });
    if (rc != MRS_OK) return rc;
    return mrs::guarded([&] {
        if (times_ms) {
            times_ms[0] = r->t_admit - r->t_submit;  // queued
            times_ms[1] = r->t_ready - r->t_admit;   // preparation
            times_ms[2] = r->t_done - r->t_submit;   // submit -> last epoch queued
        }
        if (r->state == kFailed) throw Error(r->status, r->error);
        MRS_HIP(hipSetDevice(sv->device));
        MRS_HIP(hipEventSynchronize(r->done_ev));
        if (times_ms) times_ms[2] = mrs::now_ms() - r->t_submit;  // submit -> result on the device
        if (out)
            MRS_HIP(hipMemcpy(out, r->out.p, r->out.bytes(), hipMemcpyDefault));
        (void)hipEventDestroy(r->done_ev);
        r->done_ev = nullptr;
        mrs::StreamScope scope(sv->batch_ctx->stream);
        r->out.release();
    });
}

int mrs_server_stats(mrs_server* sv, int64_t* out) {
    return mrs::guarded([&] {
        if (!sv || !out) throw Error(MRS_INVALID_ARGUMENT, "NULL argument");
        std::lock_guard<std::mutex> lk(sv->mu);
        out[0] = sv->next_id - 1;
        out[1] = sv->completed;
        out[2] = sv->failed;
        out[3] = sv->ticks;
        out[4] = sv->peak_resident;
        out[5] = (int64_t)sv->peak_in_use;
        out[6] = (int64_t)sv->budget;
    });
}

int mrs_server_destroy(mrs_server* sv) {
    return mrs::guarded([&] {
        if (!sv) return;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->stopping = true;
        }
        sv->cv_work.notify_all();
        sv->cv_batch.notify_all();
        for (auto& t : sv->workers)
            if (t.joinable()) t.join();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
        }
        sv->cv_batch.notify_all();
        if (sv->batcher.joinable()) sv->batcher.join();
        if (sv->batch_ctx) {
            mrs::StreamScope scope(sv->batch_ctx->stream);
            for (auto& kv : sv->all) {
                if (kv.second->done_ev) (void)hipEventDestroy(kv.second->done_ev);
                kv.second->st.reset();
                kv.second->out.release();
            }
            sv->all.clear();
            sv->dev_desc.release();
            (void)hipStreamSynchronize(sv->batch_ctx->stream);
        }
        for (int q = 0; q < mrs_server::kRing; q++) {
            if (sv->host_desc[q]) (void)hipHostFree(sv->host_desc[q]);
            if (sv->tick_ev[q]) (void)hipEventDestroy(sv->tick_ev[q]);
        }
        for (mrs_ctx* c : sv->prep_ctx) mrs_ctx_destroy(c);
        if (sv->batch_ctx) mrs_ctx_destroy(sv->batch_ctx);
        delete sv;
    });
}

}  // extern "C"
