// Device memory kept for reuse by one context: the logic of DevicePool,
// written against a small runtime interface RT so the same code runs on HIP
// (DevicePool, common.hip) and on a simulated multi-stream runtime in the CPU
// lifecycle harness (tests/native).
//
// A block belongs to one home stream, the stream of its context.  A block
// released on the home stream needs nothing more: everything that used it
// was queued there before the release, so stream order keeps its old and new
// users apart when the next user works on the home stream too.  A block
// released from another stream carries an event recorded on that stream at
// the release; and a block handed to a user on another stream than the one
// it was released on is fenced the same way (an event on the releasing
// stream, or on the home stream, that the user's stream waits for).  Without
// this, a block released from one stream could be reused on another while
// kernels queued on the first still read and wrote it.
//
// RT provides: types Stream and Event; static Event record(Stream) (create an
// event and record it), void wait(Stream, Event) (the stream's later work
// waits for the event), void host_wait(Event), void destroy(Event),
// void* alloc(size_t) (nullptr on failure), void release(void*),
// void sync_all() (every stream of the device drained).
#pragma once

#include <cstddef>
#include <map>
#include <mutex>

namespace mrs {

template <class RT>
class BlockCache {
  public:
    using Stream = typename RT::Stream;
    using Event = typename RT::Event;

    explicit BlockCache(Stream home) : home_(home) {}
    ~BlockCache() { trim(); }
    BlockCache(const BlockCache&) = delete;
    BlockCache& operator=(const BlockCache&) = delete;

    // A block of at least `bytes` (rounded up to 256), reused when a cached
    // one of at most twice the size exists.  `user` is the stream the caller
    // will use it on (nullptr: the host).  Returns nullptr if allocation fails
    // even after the cache was returned to the device.
    void* get(size_t bytes, size_t* got, Stream user) {
        bytes = (bytes + 255) / 256 * 256;
        Block b{};
        bool hit = false;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.lower_bound(bytes);
            if (it != free_.end() && it->first <= 2 * bytes) {
                b = it->second;
                *got = it->first;
                cached_ -= it->first;
                free_.erase(it);
                hit = true;
            }
        }
        if (hit) {
            fence(b, user);
            return b.p;
        }
        void* p = RT::alloc(bytes);
        if (!p) {
            trim();
            p = RT::alloc(bytes);
        }
        *got = bytes;
        return p;
    }

    // Return a block; `from` is the stream the caller released it on (the
    // stream of the last work that used it; nullptr: unknown, taken as home).
    void put(void* p, size_t bytes, Stream from) {
        Block b{p, from ? from : home_, Event{}, false};
        if (b.stream != home_) {
            b.ev = RT::record(b.stream);
            b.has_ev = true;
        }
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(bytes, b);
        cached_ += bytes;
    }

    // Wait for the device, then free every cached block.
    void trim() {
        std::multimap<size_t, Block> blocks;
        {
            std::lock_guard<std::mutex> lk(mu_);
            blocks.swap(free_);
            cached_ = 0;
        }
        if (blocks.empty()) return;
        RT::sync_all();  // a cached block may still be used by queued work
        for (auto& kv : blocks) {
            if (kv.second.has_ev) RT::destroy(kv.second.ev);
            RT::release(kv.second.p);
        }
    }

    size_t cached_bytes() {
        std::lock_guard<std::mutex> lk(mu_);
        return cached_;
    }

    Stream home() const { return home_; }

  private:
    struct Block {
        void* p;
        Stream stream;  // the stream it was released on
        Event ev;       // recorded there at the release (released off the home stream)
        bool has_ev;
    };

    // order the new user's work after the block's last use
    void fence(Block& b, Stream user) {
        if (user && user == b.stream) {  // same stream: stream order suffices
            if (b.has_ev) RT::destroy(b.ev);
            return;
        }
        Event ev = b.has_ev ? b.ev : RT::record(b.stream);
        if (user) RT::wait(user, ev);
        else RT::host_wait(ev);
        RT::destroy(ev);
    }

    Stream home_;
    std::mutex mu_;
    std::multimap<size_t, Block> free_;
    size_t cached_ = 0;
};

}  // namespace mrs
