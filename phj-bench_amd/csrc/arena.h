// arena.h — the best-fit device memory pool of one context.
//
// Memory comes from the backend (hipMalloc on the device) in segments of at
// least kSegmentBytes; freed blocks coalesce with free neighbours of their
// segment and are reused by later allocations of any size. Every allocation
// and every kernel of a context runs on the context's one stream, so a block
// freed there may be reused at once: stream order puts the new user's work
// after the old user's. Blocks never move to another pool (another stream).
// Segments go back to the backend only when an allocation fails, or when more
// than 60% of the device's memory sits free in this pool, and only after the
// stream's queued work (which may still use them) has finished.
//
// Ranges marked slow (mark_slow) are avoided by alloc_avoiding(), which places
// a block outside them in a free block's unmarked part when one has room (else
// it is alloc()); alloc() uses them as usual. A range is forgotten when its segment
// goes back to the backend.
//
// Backend interface (static members): Stream; void *alloc(uint64_t) (null on
// failure); void free(void *); void sync(Stream); bool mem_info(uint64_t
// &free, uint64_t &total). The header needs no HIP, so the CPU tests drive the
// pool with a recording backend.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace phjd {

// Thrown by the pool: out of memory (or over the pool's limit), or a pointer
// the pool does not own.
struct ArenaError : std::runtime_error {
    enum Kind { kOutOfMemory, kForeignPointer } kind;
    ArenaError(Kind k, const std::string &m) : std::runtime_error(m), kind(k) {}
};

template <class Dev> class BasicArena {
public:
    using Stream = typename Dev::Stream;
    static constexpr uint64_t kAlign = 256;
    static constexpr uint64_t kSegmentBytes = 1ull << 30;
    static constexpr uint64_t kMinSplit = 1ull << 20;

    ~BasicArena() { release_cached(); }

    void bind(Stream s) {
        stream_ = s;
        bound_ = true;
        uint64_t free_mem = 0, total = 0;
        if (Dev::mem_info(free_mem, total)) {
            cache_limit_ = total / 10 * 6;
            slow_limit_ = total / 8;
        }
    }

    void *alloc(uint64_t bytes) {
        const uint64_t sz = (std::max<uint64_t>(bytes, 1) + kAlign - 1) & ~(kAlign - 1);
        if (limit_ != 0 && current_ + sz > limit_)
            throw ArenaError(ArenaError::kOutOfMemory, "device allocation of " + std::to_string(sz) +
                                                           " bytes exceeds the context's memory limit");
        Block *b = take(sz);
        if (b == nullptr)
            b = grow(sz);
        if (b == nullptr) {
            release_cached();
            b = grow(sz);
        }
        if (b == nullptr)
            throw ArenaError(ArenaError::kOutOfMemory, "device allocation of " + std::to_string(sz) + " bytes failed");
        return commit(b);
    }

    // alloc() outside the slow ranges, in memory the pool already holds;
    // falls back to alloc() when no free block has such a place. It never
    // takes a new segment for the sake of avoidance: a device allocation of
    // many GB inside a caller's loop costs far more than a slow placement.
    void *alloc_avoiding(uint64_t bytes) {
        if (slow_.empty())
            return alloc(bytes);
        const uint64_t sz = (std::max<uint64_t>(bytes, 1) + kAlign - 1) & ~(kAlign - 1);
        if (limit_ != 0 && current_ + sz > limit_)
            return alloc(bytes); // throws
        Block *b = take_avoiding(sz);
        return b != nullptr ? commit(b) : alloc(bytes);
    }

    // Marks [p, p + bytes) slow. Returns false (and marks nothing) when the
    // range is already marked or the marked bytes would pass an eighth of the
    // device's memory.
    bool mark_slow(const void *p, uint64_t bytes) {
        char *a = static_cast<char *>(const_cast<void *>(p));
        for (auto &r : slow_)
            if (r.first <= a && a + bytes <= r.second)
                return false;
        if (slow_bytes_ + bytes > slow_limit_)
            return false;
        slow_.emplace_back(a, a + bytes);
        slow_bytes_ += bytes;
        return true;
    }
    uint64_t slow_bytes() const { return slow_bytes_; }
    size_t slow_ranges() const { return slow_.size(); }

    void free(void *p) {
        if (p == nullptr)
            return;
        auto it = live_.find(p);
        if (it == live_.end())
            throw ArenaError(ArenaError::kForeignPointer, "pointer not allocated by this context");
        Block *b = it->second;
        live_.erase(it);
        current_ -= b->size;
        free_bytes_ += b->size;
        b->free = true;
        b = coalesce(b);
        free_.emplace(b->size, b);
        if (free_bytes_ > cache_limit_)
            release_cached();
    }

    // free() for cleanup paths that must not throw (destructors during unwinding).
    void free_quiet(void *p) noexcept {
        try {
            free(p);
        } catch (...) {
        }
    }

    bool owns(const void *p) const { return live_.count(const_cast<void *>(p)) != 0; }

    // Returns segments with no live block to the backend, after the stream's
    // queued work has finished.
    void release_cached() {
        if (bound_)
            Dev::sync(stream_);
        for (auto it = free_.begin(); it != free_.end();) {
            Block *b = it->second;
            auto seg = segments_.find(b->ptr);
            if (seg != segments_.end() && b->prev == nullptr && b->next == nullptr && b->size == seg->second) {
                forget_slow(b->ptr, b->ptr + b->size);
                Dev::free(b->ptr);
                free_bytes_ -= b->size;
                segments_.erase(seg);
                delete b;
                it = free_.erase(it);
            } else {
                ++it;
            }
        }
    }

    uint64_t current() const { return current_; }
    uint64_t peak() const { return peak_; }
    uint64_t segments() const { return segments_.size(); }
    void reset_peak() { peak_ = current_; }
    // Upper bound on the bytes held by live blocks (0: none).
    void set_limit(uint64_t bytes) { limit_ = bytes; }

private:
    struct Block {
        char *ptr;
        uint64_t size;
        bool free;
        Block *prev, *next; // neighbours in the segment, by address
    };

    void erase_free(Block *b) {
        auto range = free_.equal_range(b->size);
        for (auto it = range.first; it != range.second; ++it)
            if (it->second == b) {
                free_.erase(it);
                return;
            }
    }

    Block *coalesce(Block *b) {
        if (b->next && b->next->free) {
            Block *n = b->next;
            erase_free(n);
            b->size += n->size;
            b->next = n->next;
            if (n->next)
                n->next->prev = b;
            delete n;
        }
        if (b->prev && b->prev->free) {
            Block *p = b->prev;
            erase_free(p);
            p->size += b->size;
            p->next = b->next;
            if (b->next)
                b->next->prev = p;
            delete b;
            b = p;
        }
        return b;
    }

    // Smallest free block of at least sz, split when the rest is worth keeping.
    Block *take(uint64_t sz) {
        auto it = free_.lower_bound(sz);
        if (it == free_.end())
            return nullptr;
        Block *b = it->second;
        free_.erase(it);
        return split_take(b, sz);
    }

    // Marks b (already out of free_) used for sz bytes; the rest stays free
    // when it is worth keeping.
    Block *split_take(Block *b, uint64_t sz) {
        if (b->size - sz >= kMinSplit) {
            Block *r = new Block{b->ptr + sz, b->size - sz, true, b, b->next};
            if (b->next)
                b->next->prev = r;
            b->next = r;
            b->size = sz;
            free_.emplace(r->size, r);
        }
        b->free = false;
        return b;
    }

    Block *grow(uint64_t sz) {
        const uint64_t seg = std::max(sz, kSegmentBytes);
        char *p = static_cast<char *>(Dev::alloc(seg));
        if (p == nullptr)
            return nullptr;
        forget_slow(p, p + seg); // an address range the backend handed out again
        segments_[p] = seg;
        free_bytes_ += seg;
        free_.emplace(seg, new Block{p, seg, true, nullptr, nullptr});
        return take(sz);
    }

    void *commit(Block *b) {
        live_[b->ptr] = b;
        free_bytes_ -= b->size;
        current_ += b->size;
        peak_ = std::max(peak_, current_);
        return b->ptr;
    }

    // Lowest aligned start in [lo, hi) of sz bytes clear of every slow range.
    char *clear_start(char *lo, char *hi, uint64_t sz) const {
        char *st = lo;
        for (bool moved = true; moved && st + sz <= hi;) {
            moved = false;
            for (auto &r : slow_)
                if (r.first < st + sz && st < r.second) {
                    const uint64_t off = static_cast<uint64_t>(r.second - lo);
                    st = lo + ((off + kAlign - 1) & ~(kAlign - 1));
                    moved = true;
                }
        }
        return st + sz <= hi ? st : nullptr;
    }

    // Smallest free block with sz bytes clear of the slow ranges; the part
    // in front of the placement stays free as a block of its own.
    Block *take_avoiding(uint64_t sz) {
        for (auto it = free_.lower_bound(sz); it != free_.end(); ++it) {
            Block *b = it->second;
            char *st = clear_start(b->ptr, b->ptr + b->size, sz);
            if (st == nullptr)
                continue;
            free_.erase(it);
            if (st == b->ptr)
                return split_take(b, sz);
            const uint64_t front = static_cast<uint64_t>(st - b->ptr);
            Block *f = new Block{b->ptr, front, true, b->prev, b};
            if (b->prev)
                b->prev->next = f;
            b->prev = f;
            b->ptr = st;
            b->size -= front;
            free_.emplace(f->size, f);
            return split_take(b, sz);
        }
        return nullptr;
    }

    void forget_slow(char *lo, char *hi) {
        for (size_t i = 0; i < slow_.size();)
            if (slow_[i].first >= lo && slow_[i].first < hi) {
                slow_bytes_ -= static_cast<uint64_t>(slow_[i].second - slow_[i].first);
                slow_[i] = slow_.back();
                slow_.pop_back();
            } else {
                ++i;
            }
    }

    Stream stream_{};
    bool bound_ = false;
    std::multimap<uint64_t, Block *> free_;
    std::unordered_map<void *, Block *> live_;
    std::unordered_map<char *, uint64_t> segments_;
    uint64_t current_ = 0, peak_ = 0, free_bytes_ = 0;
    uint64_t cache_limit_ = 64ull << 30;
    uint64_t limit_ = 0;
    std::vector<std::pair<char *, char *>> slow_; // [begin, end) address ranges
    uint64_t slow_bytes_ = 0, slow_limit_ = 16ull << 30;
};

} // namespace phjd
