// CPU lifecycle harness for the device-memory block cache (BlockCache in
// manifolds-rs_amd/csrc/block_cache.hpp, the code behind DevicePool), run
// under AddressSanitizer and ThreadSanitizer by tests/test_native_harness.py.
//
// The runtime is simulated: a stream is a thread that executes its queue of
// operations in order, an event is a flag set by the operation recorded on its
// stream, and "waiting for an event" on a stream is an operation that blocks
// until the flag is set.  Memory is plain host memory, so ASan sees every
// block's lifetime and TSan every cross-thread access.
//
// The scenario is the fit server's: prep workers, each with its own stream and
// cache, allocate a request's blocks and fill them on their stream, wait for
// that stream, and hand the request to a batcher thread; the batcher queues
// several "epochs" on its own stream that check and rewrite every block, then
// queues the output and returns the blocks to the prep worker's cache at once,
// from its own stream, without waiting.  The worker's next request can take
// those blocks while the batcher's queued work still uses them.  The cache
// must order the new owner's work after that work: a block whose contents
// change under a queued epoch is counted as corrupted, and TSan reports the
// unordered accesses as a data race.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../../manifolds-rs_amd/csrc/block_cache.hpp"

namespace {

struct EvState {
    std::mutex mu;
    std::condition_variable cv;
    bool done = false;
    void set() {
        std::lock_guard<std::mutex> lk(mu);
        done = true;
        cv.notify_all();
    }
    void wait() {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return done; });
    }
};

class SimStream {
  public:
    SimStream() : th_([this] { run(); }) {}
    ~SimStream() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
    }
    void push(std::function<void()> f) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.push_back(std::move(f));
        }
        cv_.notify_all();
    }
    void drain() {  // host waits until everything queued so far has run
        auto ev = std::make_shared<EvState>();
        push([ev] { ev->set(); });
        ev->wait();
    }

  private:
    void run() {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                f = std::move(q_.front());
                q_.pop_front();
            }
            f();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> q_;
    bool stop_ = false;
    std::thread th_;
};

std::mutex g_streams_mu;
std::vector<SimStream*> g_streams;
std::atomic<long> g_events_live{0}, g_blocks_live{0};

struct SimEvent {
    std::shared_ptr<EvState> st = std::make_shared<EvState>();
};

struct SimRT {
    using Stream = SimStream*;
    using Event = SimEvent*;
    static Event record(Stream s) {
        auto* e = new SimEvent();
        g_events_live++;
        s->push([st = e->st] { st->set(); });
        return e;
    }
    static void wait(Stream s, Event e) {
        s->push([st = e->st] { st->wait(); });
    }
    static void host_wait(Event e) { e->st->wait(); }
    static void destroy(Event e) {
        g_events_live--;
        delete e;
    }
    static void* alloc(size_t bytes) {
        g_blocks_live++;
        return std::malloc(bytes);
    }
    static void release(void* p) {
        g_blocks_live--;
        std::free(p);
    }
    static void sync_all() {
        std::lock_guard<std::mutex> lk(g_streams_mu);
        for (SimStream* s : g_streams) s->drain();
    }
};

using Cache = mrs::BlockCache<SimRT>;

struct Block {
    uint32_t* p;
    size_t words, cap;
};

struct Request {
    uint32_t id;
    Cache* cache;
    std::vector<Block> blocks;
};

std::atomic<long> g_corrupted{0}, g_checked{0};

void check_fill(const Block& b, uint32_t id) {
    for (size_t w = 0; w < b.words; w++)
        if (b.p[w] != id) {
            g_corrupted++;
            return;
        }
    g_checked++;
}

}  // namespace

int main(int argc, char** argv) {
    const int n_requests = argc > 1 ? std::atoi(argv[1]) : 300;
    const int n_workers = 2, epochs = 3;
    SimStream batch_stream;
    std::vector<std::unique_ptr<SimStream>> prep_streams;
    std::vector<std::unique_ptr<Cache>> caches;
    for (int w = 0; w < n_workers; w++) {
        prep_streams.emplace_back(new SimStream());
        caches.emplace_back(new Cache(prep_streams.back().get()));
    }
    {
        std::lock_guard<std::mutex> lk(g_streams_mu);
        g_streams.push_back(&batch_stream);
        for (auto& s : prep_streams) g_streams.push_back(s.get());
    }

    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request*> ready;
    std::atomic<int> next{0};
    int workers_left = n_workers;

    auto worker = [&](int w) {
        std::mt19937 rng(1234 + w);
        SimStream* s = prep_streams[w].get();
        Cache* cache = caches[w].get();
        for (;;) {
            const int r = next++;
            if (r >= n_requests) break;
            auto* req = new Request{(uint32_t)r + 1, cache, {}};
            const int nb = 2 + (int)(rng() % 3);
            for (int q = 0; q < nb; q++) {
                const size_t words = (size_t)256 << (rng() % 5);
                Block b{};
                b.words = words;
                // one request in eight takes a block for host use (no stream)
                const bool host = rng() % 8 == 0;
                b.p = static_cast<uint32_t*>(cache->get(words * 4, &b.cap, host ? nullptr : s));
                if (host) {
                    for (size_t x = 0; x < words; x++) b.p[x] = req->id;
                } else {
                    s->push([b, id = req->id] {
                        for (size_t x = 0; x < b.words; x++) b.p[x] = id;
                    });
                }
                req->blocks.push_back(b);
            }
            s->drain();  // the prep worker waits for its stream before the hand-over
            {
                std::lock_guard<std::mutex> lk(mu);
                ready.push_back(req);
            }
            cv.notify_all();
        }
        std::lock_guard<std::mutex> lk(mu);
        workers_left--;
        cv.notify_all();
    };

    auto batcher = [&] {
        std::vector<std::pair<Request*, int>> active;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !ready.empty() || !active.empty() || workers_left == 0; });
                while (!ready.empty()) {
                    active.push_back({ready.front(), 0});
                    ready.pop_front();
                }
                if (active.empty() && workers_left == 0) break;
            }
            // one tick: every resident request advances by one epoch
            std::vector<std::pair<Request*, int>> keep;
            for (auto& a : active) {
                Request* req = a.first;
                batch_stream.push([blocks = req->blocks, id = req->id] {
                    std::this_thread::sleep_for(std::chrono::microseconds(30));
                    for (const Block& b : blocks) {
                        check_fill(b, id);
                        for (size_t x = 0; x < b.words; x++) b.p[x] = id;  // the epoch's writes
                    }
                });
                if (++a.second < epochs) {
                    keep.push_back(a);
                    continue;
                }
                // retire: queue the output, return the blocks at once from this stream
                batch_stream.push([blocks = req->blocks, id = req->id] {
                    std::this_thread::sleep_for(std::chrono::microseconds(50));
                    for (const Block& b : blocks) check_fill(b, id);
                });
                for (const Block& b : req->blocks) req->cache->put(b.p, b.cap, &batch_stream);
                delete req;
            }
            active.swap(keep);
        }
        batch_stream.drain();
    };

    std::vector<std::thread> th;
    for (int w = 0; w < n_workers; w++) th.emplace_back(worker, w);
    std::thread bt(batcher);
    for (auto& t : th) t.join();
    bt.join();
    for (auto& c : caches) c->trim();
    caches.clear();
    {
        std::lock_guard<std::mutex> lk(g_streams_mu);
        g_streams.clear();
    }
    std::printf("requests %d: %ld block checks, %ld corrupted; events live %ld, blocks live %ld\n",
                n_requests, g_checked.load(), g_corrupted.load(), g_events_live.load(),
                g_blocks_live.load());
    const bool ok = g_corrupted == 0 && g_events_live == 0 && g_blocks_live == 0 && g_checked > 0;
    std::printf("%s\n", ok ? "HARNESS_OK" : "HARNESS_FAIL");
    return ok ? 0 : 1;
}
