// arena_selftest — CPU test of the device pool's reuse rules (arena.h), driven
// by a backend that hands out fake device addresses and records every call.
// Built and run by tests/test_host_units.py; prints "ok" and exits 0 when
// every rule holds.
//
// Rules checked:
//   1. a freed block is reused by the next allocation that fits, from the
//      same segment, without a backend allocation;
//   2. freed neighbours coalesce, so a larger block fits where two were;
//   3. two pools never hand out each other's blocks (blocks never move to
//      another stream): their segments are distinct;
//   4. segments go back to the backend only when an allocation fails or the
//      free bytes pass the cache limit, and only after the pool's stream
//      has been synchronised;
//   5. a memory limit refuses an allocation without changing what is held;
//   6. freeing a pointer the pool does not own is refused;
//   7. alloc_avoiding() places a block clear of the ranges marked slow in
//      the unmarked part of a free block, takes no new segment for that (it
//      is alloc() when no free block has room), alloc() still uses the
//      ranges, and a range is forgotten with its segment.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "arena.h"

namespace {

struct Log {
    std::vector<std::string> events;
    uint64_t next_addr = 1ull << 40;
    uint64_t device_total = 100ull << 30; // 100 GiB "device"
    uint64_t in_use = 0;
    int fail_next_allocs = 0;
} g;

struct FakeDev {
    using Stream = int;
    static void *alloc(uint64_t bytes) {
        if (g.fail_next_allocs > 0 || g.in_use + bytes > g.device_total) {
            if (g.fail_next_allocs > 0)
                --g.fail_next_allocs;
            g.events.push_back("alloc-fail");
            return nullptr;
        }
        void *p = reinterpret_cast<void *>(g.next_addr);
        g.next_addr += bytes + (1ull << 30);
        g.in_use += bytes;
        g.events.push_back("alloc " + std::to_string(bytes >> 20) + "M");
        return p;
    }
    static void free(void *) { g.events.push_back("free"); }
    static void sync(int s) { g.events.push_back("sync " + std::to_string(s)); }
    static bool mem_info(uint64_t &f, uint64_t &t) {
        f = g.device_total - g.in_use;
        t = g.device_total;
        return true;
    }
};

using Pool = phjd::BasicArena<FakeDev>;
int failures = 0;

void expect(bool ok, const char *what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}

size_t count(const std::string &prefix) {
    size_t n = 0;
    for (const auto &e : g.events)
        n += e.compare(0, prefix.size(), prefix) == 0;
    return n;
}

} // namespace

int main() {
    constexpr uint64_t MiB = 1ull << 20;
    {
        Pool a;
        a.bind(7);
        // rule 1
        void *x = a.alloc(64 * MiB);
        expect(count("alloc ") == 1 && a.segments() == 1, "first allocation takes one 1 GiB segment");
        a.free(x);
        void *y = a.alloc(48 * MiB);
        expect(y == x, "a freed block is reused by the next allocation that fits");
        expect(count("alloc ") == 1, "reuse needs no backend allocation");
        // rule 2
        void *p = a.alloc(100 * MiB), *q = a.alloc(100 * MiB), *r = a.alloc(100 * MiB);
        a.free(p);
        a.free(q);
        void *big = a.alloc(200 * MiB);
        expect(big == p, "freed neighbours coalesce into one block");
        expect(a.current() == (48 + 200 + 100) * MiB, "bytes held are the live blocks");
        // rule 3
        Pool b;
        b.bind(9);
        void *z = b.alloc(10 * MiB);
        expect(b.segments() == 1 && count("alloc ") == 2, "a second pool takes a segment of its own");
        a.free(r);
        void *w = b.alloc(100 * MiB);
        expect(w != r, "a block freed in one pool is never handed out by another");
        b.free(w);
        b.free(z);
        // rule 4: nothing returned while the pool stays under the cache limit
        expect(count("free") == 0, "freed segments stay cached");
        const size_t syncs_before = count("sync 7");
        g.fail_next_allocs = 1; // the next backend allocation fails once
        void *huge = a.alloc(3ull << 30);
        expect(huge != nullptr, "a failed backend allocation is retried after releasing cached segments");
        expect(count("sync 7") == syncs_before + 1, "the pool's own stream is synchronised before segments go back");
        const size_t sync_at = [&] {
            for (size_t i = 0; i < g.events.size(); ++i)
                if (g.events[i] == "sync 7")
                    return i;
            return g.events.size();
        }();
        expect(sync_at < g.events.size() && count("free") == 0,
               "segments still holding live blocks are not returned");
        a.free(huge);
        // rule 5
        a.set_limit(a.current() + 16 * MiB);
        const uint64_t held = a.current();
        bool refused = false;
        try {
            a.alloc(32 * MiB);
        } catch (const phjd::ArenaError &e) {
            refused = e.kind == phjd::ArenaError::kOutOfMemory;
        }
        expect(refused && a.current() == held, "the memory limit refuses without changing what is held");
        a.set_limit(0);
        // rule 6
        bool foreign = false;
        try {
            a.free(z);
        } catch (const phjd::ArenaError &e) {
            foreign = e.kind == phjd::ArenaError::kForeignPointer;
        }
        expect(foreign, "freeing another pool's pointer is refused");
        a.free(y);
        a.free(big);
        // cache limit: 60% of the 100 GiB device; 70 GiB freed at once goes back
        void *hold = a.alloc(70ull << 30);
        const size_t frees = count("free");
        a.free(hold);
        expect(count("free") > frees, "free bytes past the cache limit go back to the backend");
    }
    {
        // rule 7
        Pool c;
        c.bind(9);
        char *seg = static_cast<char *>(c.alloc(1024 * MiB)); // one 1 GiB segment
        c.free(seg);
        expect(c.mark_slow(seg + 100 * MiB, 200 * MiB), "a range is marked slow");
        expect(!c.mark_slow(seg + 150 * MiB, 10 * MiB), "a range inside a marked one is not marked again");
        const size_t allocs = count("alloc ");
        char *p = static_cast<char *>(c.alloc_avoiding(50 * MiB));
        expect(p == seg, "a block that fits in front of the slow range starts the segment");
        char *q = static_cast<char *>(c.alloc_avoiding(150 * MiB));
        expect(q == seg + 300 * MiB, "a block that does not fit in front goes past the slow range");
        expect(count("alloc ") == allocs, "both fit the cached segment without a backend allocation");
        char *r = static_cast<char *>(c.alloc_avoiding(100 * MiB));
        expect(r == seg + 450 * MiB, "the rest of the segment past the slow range is used");
        char *rest = static_cast<char *>(c.alloc(474 * MiB)); // the segment's tail
        char *t = static_cast<char *>(c.alloc_avoiding(200 * MiB));
        expect(rest == seg + 550 * MiB && t == seg + 50 * MiB && count("alloc ") == allocs,
               "no clear place left: an ordinary allocation over the slow range, without a new segment");
        c.free(t);
        c.free(rest);
        char *u = static_cast<char *>(c.alloc(40 * MiB));
        expect(u >= seg + 50 * MiB && u < seg + 300 * MiB, "alloc() uses the slow range as usual");
        for (char *x : {p, q, r, u})
            c.free(x);
        expect(c.slow_ranges() == 1, "the range stays marked while its segment is held");
        c.release_cached();
        expect(c.slow_ranges() == 0 && c.slow_bytes() == 0, "a range is forgotten with its segment");
        char *v = static_cast<char *>(c.alloc_avoiding(64 * MiB));
        expect(v != nullptr, "with no range marked, alloc_avoiding() is alloc()");
        c.free(v);
    }
    expect(count("free") == count("alloc "), "every segment is returned when the pools are destroyed");
    if (failures == 0)
        std::printf("ok\n");
    return failures == 0 ? 0 : 1;
}
