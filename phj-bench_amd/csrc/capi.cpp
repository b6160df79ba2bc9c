// capi.cpp — the C ABI of libphj_hip.so: context, device arena, relation
// staging, and the host orchestration of the three schemes.
//
// Scheme orchestration follows the reference's phase structure:
//   runPHJ    (src/PHJ.cpp:19-150):    build shuffle -> probe shuffle -> per-leaf build+probe
//   runCHJ    (src/CHJ.cpp:16-119):    sub-table build -> probe
//   runPhjBep (src/PHJBep.cpp:113-790): build shuffle -> probe side in slices whose two
//                                       pass buffers fit the budget, each slice
//                                       partitioned to the leaves and joined at once
// Every kernel is launched on the context's one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/phj_gpu.h"
#include "arena.h"
#include "gen.h"
#include "kernels.h"

using namespace phjd;

namespace {

struct PhjError : std::runtime_error {
    phj_status code;
    PhjError(phj_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            throw PhjError(e_ == hipErrorOutOfMemory ? PHJ_ERR_OUT_OF_MEMORY : PHJ_ERR_HIP,                           \
                           std::string(#expr) + ": " + hipGetErrorString(e_));                                         \
    } while (0)

void invalid(const std::string &m) { throw PhjError(PHJ_ERR_INVALID_ARGUMENT, m); }

// ---------------------------------------------------------------- device memory
// The context's pool (arena.h) over hipMalloc, with the context stream.
struct HipBackend {
    using Stream = hipStream_t;
    static void *alloc(uint64_t bytes) {
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return p;
    }
    static void free(void *p) { (void)hipFree(p); }
    static void sync(hipStream_t s) {
        if (s != nullptr)
            (void)hipStreamSynchronize(s);
    }
    static bool mem_info(uint64_t &free_mem, uint64_t &total) {
        size_t f = 0, t = 0;
        if (hipMemGetInfo(&f, &t) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        free_mem = f;
        total = t;
        return true;
    }
};
using Arena = BasicArena<HipBackend>;

// ---------------------------------------------------------------- kernel timer
class EventTimer : public KernelTimer {
public:
    ~EventTimer() override {
        for (auto e : pool_)
            (void)hipEventDestroy(e);
    }
    void begin(const char *name, hipStream_t s) override {
        hipEvent_t a = get();
        (void)hipEventRecord(a, s);
        open_.push_back({name, a, nullptr});
    }
    void end(hipStream_t s) override {
        hipEvent_t b = get();
        (void)hipEventRecord(b, s);
        open_.back().b = b;
        pending_.push_back(open_.back());
        open_.pop_back();
    }
    // After the stream has synchronised.
    void resolve() {
        for (auto &p : pending_) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                auto &t = totals_[p.name];
                t.first += ms;
                t.second += 1;
            }
            free_.push_back(p.a);
            free_.push_back(p.b);
        }
        pending_.clear();
    }
    void reset() { totals_.clear(); }
    bool get_total(const std::string &n, double &ms, uint64_t &cnt) const {
        auto it = totals_.find(n);
        if (it == totals_.end())
            return false;
        ms = it->second.first;
        cnt = it->second.second;
        return true;
    }
    std::string names() const {
        std::string s;
        for (auto &kv : totals_) {
            if (!s.empty())
                s += ",";
            s += kv.first;
        }
        return s;
    }

private:
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    hipEvent_t get() {
        if (!free_.empty()) {
            hipEvent_t e = free_.back();
            free_.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        pool_.push_back(e);
        return e;
    }
    std::vector<hipEvent_t> pool_, free_;
    std::vector<Rec> open_, pending_;
    std::map<std::string, std::pair<double, uint64_t>> totals_;
};

} // namespace

// ---------------------------------------------------------------- allocation-free scratch
// Device memory of phj_hashtable_build_find_noalloc: the hash-table cells, the
// build rows' chain successors and one chunk of probe-row results, in one
// file-scope array that every GPU holds once for the life of the process.
// Host code never passes `scratch` itself to a kernel (on the host the name
// is only a stand-in): it uses the address hipGetSymbolAddress returns for
// the current device, looked up once per device when its first context is
// created.
constexpr uint64_t kNoAllocMaxBuild = 1ull << 20;              // build rows
constexpr uint64_t kNoAllocMaxCells = 2 * kNoAllocMaxBuild;    // bit_ceil(2 * rows), src/HashTable.h:72-77
constexpr uint64_t kNoAllocProbeChunk = 1ull << 20;            // probe rows per lookup launch
constexpr uint64_t kNoAllocCellsOff = 0;
constexpr uint64_t kNoAllocNextOff = kNoAllocCellsOff + kNoAllocMaxCells * 16;
constexpr uint64_t kNoAllocRowsOff = kNoAllocNextOff + kNoAllocMaxBuild * 4;
constexpr uint64_t kNoAllocWordsOff = kNoAllocRowsOff + kNoAllocProbeChunk * 4; // cell offsets, side head
constexpr uint64_t kNoAllocBytes = kNoAllocWordsOff + 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (current), looked up once.
hipError_t scratch_address(int device, void **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices)
        return hipErrorInvalidDevice;
    if (g_scratch_ptr[device] == nullptr) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}
} // namespace

constexpr uint32_t kBatchClasses = 33;           // size classes floor(log2(build rows)), 0..32
constexpr uint32_t kBatchDefaultLeafRows = 2048; // leaf target of every class until tuned

namespace {
// Pinned host memory for the small arrays a call uploads (offsets, bit
// windows, segment lists) and the values it reads back, so that no copy goes
// through pageable memory (which can hold the host until the stream has
// caught up). A bump allocator over blocks kept for the context's life;
// reset() only once the stream has finished with everything handed out.
class PinnedArena {
public:
    ~PinnedArena() {
        for (auto &b : blocks_)
            (void)hipHostFree(b.p);
    }
    template <class T> T *get(uint64_t n) { return static_cast<T *>(bytes(std::max<uint64_t>(n, 1) * sizeof(T))); }
    void *bytes(uint64_t sz) {
        sz = (sz + 63) & ~uint64_t{63};
        while (cur_ < blocks_.size() && off_ + sz > blocks_[cur_].cap) {
            ++cur_;
            off_ = 0;
        }
        if (cur_ == blocks_.size()) {
            Block b{nullptr, std::max<uint64_t>(sz, 8ull << 20)};
            if (hipHostMalloc(reinterpret_cast<void **>(&b.p), b.cap, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                throw std::bad_alloc();
            }
            blocks_.push_back(b);
            off_ = 0;
        }
        void *p = blocks_[cur_].p + off_;
        off_ += sz;
        return p;
    }
    void reset() {
        cur_ = 0;
        off_ = 0;
    }

private:
    struct Block {
        char *p;
        uint64_t cap;
    };
    std::vector<Block> blocks_;
    size_t cur_ = 0;
    uint64_t off_ = 0;
};
} // namespace

struct phj_ctx {
    int device = 0;
    void *scratch = nullptr; // this device's `scratch` (device address)
    hipStream_t stream = nullptr;
    std::string err;
    Arena arena;
    EventTimer timer;
    bool timing = false;
    std::unordered_map<uint64_t, std::vector<void *>> owned; // handle -> buffers
    // handle -> buffers shared by the results of one batch group (freed with the last)
    std::unordered_map<uint64_t, std::shared_ptr<std::vector<void *>>> shared;
    uint64_t next_handle = 1;
    // Host-visible words a join reads back without draining the stream: a
    // value copied there is waited for through `note_event` only.
    uint32_t *pinned = nullptr;
    hipEvent_t note_event = nullptr;
    PinnedArena pinned_io; // uploads and readbacks of batched joins
    // Leaf size (build rows) batched joins aim for, by size class
    // floor(log2(build rows)); set by phj_batch_autotune.
    uint32_t batch_leaf_rows[kBatchClasses];
    phj_ctx() { std::fill(std::begin(batch_leaf_rows), std::end(batch_leaf_rows), kBatchDefaultLeafRows); }

    // Placement watch of the radix scatter. The same scatter runs up to 30%
    // slower into some placements of its output than into others, reproducibly
    // for a given block within a process (measured on MI355X at 2^30 rows; the
    // cause is not in the kernel). Passes over at least kWatchRows rows are
    // timed; one slower than the fastest pass of its shape by more than 15%
    // marks its largest output block slow in the pool, and later pass outputs
    // are placed clear of the marked blocks.
    struct PassWatch {
        uint64_t shape = 0;
        hipEvent_t a = nullptr, b = nullptr;
        void *block = nullptr; // the pass's largest output block
        uint64_t bytes = 0;
        float ms = -1;
    };
    static constexpr uint64_t kWatchRows = 1ull << 26;
    static constexpr size_t kWatchKeep = 64;
    std::vector<PassWatch> watch_open, watch_done;
    std::vector<hipEvent_t> watch_events; // unused events, for reuse
    std::unordered_map<uint64_t, float> watch_best;
    uint64_t watch_marked = 0;

    hipEvent_t watch_event() {
        if (!watch_events.empty()) {
            hipEvent_t e = watch_events.back();
            watch_events.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return e;
    }
    // Reads the finished timed passes (never waits) and marks slow blocks.
    void watch_review() {
        for (size_t i = 0; i < watch_open.size();) {
            PassWatch &w = watch_open[i];
            if (hipEventQuery(w.b) != hipSuccess) {
                (void)hipGetLastError();
                ++i;
                continue;
            }
            if (hipEventElapsedTime(&w.ms, w.a, w.b) != hipSuccess) {
                (void)hipGetLastError();
                w.ms = -1;
            }
            watch_events.push_back(w.a);
            watch_events.push_back(w.b);
            if (w.ms > 0) {
                auto it = watch_best.find(w.shape);
                if (it == watch_best.end() || w.ms < it->second)
                    watch_best[w.shape] = w.ms;
                if (watch_done.size() == kWatchKeep)
                    watch_done.erase(watch_done.begin());
                watch_done.push_back(w);
            }
            watch_open[i] = watch_open.back();
            watch_open.pop_back();
        }
        for (auto &w : watch_done)
            if (w.block != nullptr && w.ms > 1.15f * watch_best[w.shape]) {
                if (arena.mark_slow(w.block, w.bytes))
                    ++watch_marked;
                w.block = nullptr;
            }
    }
    void watch_release() {
        for (auto &w : watch_open) {
            (void)hipEventDestroy(w.a);
            (void)hipEventDestroy(w.b);
        }
        for (auto e : watch_events)
            (void)hipEventDestroy(e);
        watch_open.clear();
        watch_events.clear();
    }

    KernelTimer *tm() { return timing ? &timer : nullptr; }
    void sync() {
        HIPCHK(hipGetLastError()); // launch failures (bad configuration, LDS size)
        HIPCHK(hipStreamSynchronize(stream));
        if (timing)
            timer.resolve();
    }
    template <class T> T *alloc(uint64_t n) { return static_cast<T *>(arena.alloc(n * sizeof(T))); }
};

namespace {

// Buffers freed when the scope ends (back to the context's pool).
struct Scratch {
    phj_ctx *ctx;
    std::vector<void *> bufs;
    explicit Scratch(phj_ctx *c) : ctx(c) {}
    ~Scratch() {
        for (void *p : bufs)
            ctx->arena.free_quiet(p);
    }
    template <class T> T *get(uint64_t n) {
        T *p = ctx->alloc<T>(std::max<uint64_t>(n, 1));
        bufs.push_back(p);
        return p;
    }
    void release(void *p) {
        auto it = std::find(bufs.begin(), bufs.end(), p);
        if (it != bufs.end()) {
            ctx->arena.free(p);
            bufs.erase(it);
        }
    }
};

// Device blocks one call holds beyond its scope's Scratch (pass outputs that
// outlive the function making them, tables kept across slices): returned to
// the context's pool when the call unwinds, unless their ownership has moved
// to a result first (clear()).
struct Held : std::vector<void *> {
    phj_ctx *ctx;
    explicit Held(phj_ctx *c) : ctx(c) {}
    Held(const Held &) = delete;
    Held &operator=(const Held &) = delete;
    ~Held() { free_all(); }
    void free_all() noexcept {
        for (void *p : *this)
            ctx->arena.free_quiet(p);
        clear();
    }
};

uint32_t tsize(uint32_t t) {
    if (t > 4)
        invalid("payload type code must be 0..4");
    return type_size(t);
}

// A relation resident on the device (copied from the host when needed).
struct DevRel {
    uint64_t rows = 0;
    const uint64_t *keys = nullptr;
    ColSet cols{};
};

DevRel stage_relation(phj_ctx *ctx, Scratch &scr, const phj_relation &r) {
    if (r.n_cols > PHJ_MAX_COLS)
        invalid("at most 8 payload columns");
    if (r.rows > 0xFFFFFFFEull)
        invalid("relations are limited to 2^32-2 rows");
    if (r.rows > 0 && r.keys == nullptr)
        invalid("null key column");
    DevRel d;
    d.rows = r.rows;
    d.cols.n = r.n_cols;
    for (uint32_t c = 0; c < r.n_cols; ++c) {
        d.cols.types[c] = r.cols[c].type;
        tsize(r.cols[c].type);
        if (r.rows > 0 && r.cols[c].data == nullptr)
            invalid("null payload column");
    }
    if (r.on_device) {
        d.keys = r.keys;
        for (uint32_t c = 0; c < r.n_cols; ++c)
            d.cols.ptr[c] = r.cols[c].data;
        return d;
    }
    uint64_t *k = scr.get<uint64_t>(r.rows);
    if (r.rows)
        HIPCHK(hipMemcpyAsync(k, r.keys, r.rows * 8, hipMemcpyHostToDevice, ctx->stream));
    d.keys = k;
    for (uint32_t c = 0; c < r.n_cols; ++c) {
        const uint64_t bytes = r.rows * tsize(r.cols[c].type);
        void *p = scr.get<uint8_t>(bytes);
        if (bytes)
            HIPCHK(hipMemcpyAsync(p, r.cols[c].data, bytes, hipMemcpyHostToDevice, ctx->stream));
        d.cols.ptr[c] = p;
    }
    return d;
}

// src/RadixPartition.cpp:554-560 and src/CLI.cpp:126,350-355.
uint32_t validate_pass_bits(const uint8_t *pass_bits, uint32_t n) {
    if (n == 0 || pass_bits == nullptr)
        invalid("radixShuffle: at least one pass required");
    if (n > PHJ_MAX_PASSES)
        invalid("at most 8 passes");
    uint32_t total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (pass_bits[i] == 0 || pass_bits[i] > 16)
            invalid("partition bits per pass must be in [1, 16]");
        total += pass_bits[i];
    }
    if (total == 0 || total > 32)
        invalid("radixShuffle: total partition bits must be in [1, 32]");
    return total;
}

// Device pass plan: the leaves depend only on the total bit count (the top
// `total` hash bits, consumed top-down), so the device splits them into the
// fewest passes of at most kDeviceMaxPassBits bits, each as even as possible.
std::vector<uint32_t> device_passes(uint32_t total) {
    if (total > 24)
        throw PhjError(PHJ_ERR_UNSUPPORTED, "more than 24 partition bits (16M leaves) is not supported on the device");
    const uint32_t n = (total + kDeviceMaxPassBits - 1) / kDeviceMaxPassBits;
    std::vector<uint32_t> v;
    for (uint32_t p = 0; p < n; ++p)
        v.push_back(total / n + (p < total % n ? 1 : 0));
    return v;
}

struct Partitioned {
    uint64_t rows = 0;
    uint64_t *keys = nullptr;
    OutColSet cols{};
    uint32_t *off = nullptr; // parts + 1
    uint64_t parts = 0;
    int64_t max_part = -1;   // rows of the largest partition when known on the host, else -1
};

// Starts reading the largest partition of P back into the context's pinned
// word; the copy completes in stream order, so work queued after it (the
// probe side's passes) keeps the device busy while the host later waits for
// this one value alone (max_part_wait).
void max_part_start(phj_ctx *ctx, const Partitioned &P, Scratch &scr) {
    uint32_t *d = scr.get<uint32_t>(1);
    leaf_max_rows(ctx->stream, P.off, static_cast<uint32_t>(P.parts), d);
    HIPCHK(hipMemcpyAsync(ctx->pinned, d, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipEventRecord(ctx->note_event, ctx->stream));
}

void max_part_wait(phj_ctx *ctx, Partitioned &P) {
    HIPCHK(hipEventSynchronize(ctx->note_event));
    P.max_part = *ctx->pinned;
}

OutColSet alloc_cols(phj_ctx *ctx, const ColSet &like, uint64_t rows, std::vector<void *> &track,
                     bool avoid_slow = false) {
    OutColSet o{};
    o.n = like.n;
    for (uint32_t c = 0; c < like.n; ++c) {
        o.types[c] = like.types[c] == kTypeRowIndex ? 2u : like.types[c]; // synthesised row indices land as u32
        const uint64_t bytes = std::max<uint64_t>(rows * type_size(o.types[c]), 1);
        o.ptr[c] = avoid_slow ? ctx->arena.alloc_avoiding(bytes) : ctx->arena.alloc(bytes);
        track.push_back(o.ptr[c]);
    }
    return o;
}

ColSet as_in(const OutColSet &o) {
    ColSet c{};
    c.n = o.n;
    for (uint32_t i = 0; i < o.n; ++i) {
        c.types[i] = o.types[i];
        c.ptr[i] = o.ptr[i];
    }
    return c;
}

// Multi-pass radix partition of `in` (segments `in_off` with `n_seg` entries,
// or the whole relation when in_off == nullptr) by `passes` starting at bit
// position `consumed` from the top. Result buffers are tracked in `keep`.
Partitioned radix_partition(phj_ctx *ctx, const DevRel &in, const uint32_t *in_off, uint32_t n_seg, uint32_t consumed,
                            const std::vector<uint32_t> &passes, std::vector<void *> &keep, const char *,
                            uint32_t flags = kRadixNtLoads) {
    Scratch scr(ctx);
    const uint64_t rows = in.rows;
    uint32_t *seg = in_off ? const_cast<uint32_t *>(in_off) : nullptr;
    if (!seg) {
        seg = scr.get<uint32_t>(2);
        launch_seg_single(ctx->stream, seg, static_cast<uint32_t>(rows));
        n_seg = 1;
    }
    const uint64_t *cur_keys = in.keys;
    ColSet cur_cols = in.cols;
    Held prev_bufs(ctx); // buffers of the previous pass output
    Partitioned out;
    out.rows = rows;
    uint32_t *cur_seg = seg;
    uint32_t cur_nseg = n_seg;
    // Between two passes the scatter leaves the next pass's digit of every row
    // in a byte column, which the next histogram reads instead of the keys.
    Held dig_prev(ctx);
    const uint8_t *dig_in = nullptr;
    for (size_t p = 0; p < passes.size(); ++p) {
        const uint32_t bits = passes[p];
        consumed += bits;
        const uint32_t shift = 64 - consumed;
        const uint64_t B = 1ull << bits;
        const uint64_t out_nseg = static_cast<uint64_t>(cur_nseg) * B;
        PassDigits dig{dig_in, nullptr, 0, 0};
        Held dig_next(ctx);
        if (p + 1 < passes.size() && passes[p + 1] <= 8 && rows > 0 && !(flags & kRadixHashEveryPass)) {
            dig.out = ctx->alloc<uint8_t>(rows);
            dig_next.push_back(dig.out);
            dig.next_shift = 64 - consumed - passes[p + 1];
            dig.next_bits = passes[p + 1];
        }
        const bool watch = rows >= phj_ctx::kWatchRows && ctx->watch_open.size() < phj_ctx::kWatchKeep;
        if (watch)
            ctx->watch_review();
        Held bufs(ctx);
        uint64_t *nkeys = static_cast<uint64_t *>(watch ? ctx->arena.alloc_avoiding(rows * sizeof(uint64_t))
                                                        : ctx->arena.alloc(std::max<uint64_t>(rows, 1) * sizeof(uint64_t)));
        bufs.push_back(nkeys);
        OutColSet ncols = alloc_cols(ctx, cur_cols, rows, bufs, watch);
        uint32_t *nseg = ctx->alloc<uint32_t>(out_nseg + 1);
        bufs.push_back(nseg);
        const uint64_t max_tiles = radix_max_tiles(rows, cur_nseg);
        RadixPassScratch rs;
        rs.seg_tile_first = scr.get<uint32_t>(2 * (static_cast<uint64_t>(cur_nseg) + 1));
        rs.hist = scr.get<uint32_t>(max_tiles * B);
        rs.block_sums = scr.get<uint32_t>(radix_max_chunks(rows, cur_nseg) * B);
        phj_ctx::PassWatch w;
        if (watch) {
            uint64_t row_bytes = sizeof(uint64_t);
            w.block = nkeys;
            w.bytes = rows * sizeof(uint64_t);
            for (uint32_t c = 0; c < ncols.n; ++c) {
                const uint64_t cb = rows * type_size(ncols.types[c]);
                row_bytes += type_size(ncols.types[c]);
                if (cb > w.bytes) {
                    w.block = ncols.ptr[c];
                    w.bytes = cb;
                }
            }
            w.shape = rows * 0x9E3779B97F4A7C15ull ^ row_bytes << 40 ^ static_cast<uint64_t>(bits) << 50 ^
                      static_cast<uint64_t>(p) << 56 ^ static_cast<uint64_t>(cur_nseg > 1) << 62;
            w.a = ctx->watch_event();
            w.b = ctx->watch_event();
            if (w.a == nullptr || w.b == nullptr || hipEventRecord(w.a, ctx->stream) != hipSuccess) {
                (void)hipGetLastError();
                if (w.a)
                    ctx->watch_events.push_back(w.a);
                if (w.b)
                    ctx->watch_events.push_back(w.b);
                w.a = w.b = nullptr;
            }
        }
        try {
            radix_pass(ctx->stream, cur_keys, cur_cols, nkeys, ncols, cur_seg, cur_nseg, rows, uniform_bits(shift, bits), nseg,
                       rs, flags, ctx->tm(), dig);
        } catch (...) {
            if (w.a != nullptr) {
                ctx->watch_events.push_back(w.a);
                ctx->watch_events.push_back(w.b);
            }
            throw;
        }
        if (w.a != nullptr) {
            if (hipEventRecord(w.b, ctx->stream) == hipSuccess) {
                ctx->watch_open.push_back(w);
            } else {
                (void)hipGetLastError();
                ctx->watch_events.push_back(w.a);
                ctx->watch_events.push_back(w.b);
            }
        }
        // previous pass output is dead once this pass has consumed it
        prev_bufs.free_all();
        prev_bufs.swap(bufs);
        dig_prev.free_all();
        dig_prev.swap(dig_next);
        dig_in = dig.out;
        cur_keys = nkeys;
        cur_cols = as_in(ncols);
        cur_seg = nseg;
        cur_nseg = static_cast<uint32_t>(out_nseg);
        out.keys = nkeys;
        out.cols = ncols;
        out.off = nseg;
        out.parts = out_nseg;
    }
    for (void *b : prev_bufs)
        keep.push_back(b);
    prev_bufs.clear();
    return out;
}

// Output buffers of a join result (capacity rows). They own their blocks:
// an unwinding call returns them to the context's pool; own_result() moves
// them to a result handle.
struct OutBuffers {
    phj_ctx *ctx = nullptr;
    uint64_t cap = 0;
    uint64_t *keys = nullptr;
    OutColSet left{}, right{};
    std::vector<void *> bufs;

    OutBuffers() = default;
    OutBuffers(const OutBuffers &) = delete;
    OutBuffers &operator=(const OutBuffers &) = delete;
    OutBuffers(OutBuffers &&o) noexcept { take(o); }
    OutBuffers &operator=(OutBuffers &&o) noexcept {
        if (this != &o) {
            release();
            take(o);
        }
        return *this;
    }
    ~OutBuffers() { release(); }
    void release() noexcept {
        for (void *p : bufs)
            ctx->arena.free_quiet(p);
        bufs.clear();
    }

private:
    void take(OutBuffers &o) noexcept {
        ctx = o.ctx;
        cap = o.cap;
        keys = o.keys;
        left = o.left;
        right = o.right;
        bufs = std::move(o.bufs);
        o.bufs.clear();
    }
};

OutBuffers alloc_output(phj_ctx *ctx, const ColSet &l, const ColSet &r, uint64_t cap) {
    OutBuffers o;
    o.ctx = ctx;
    o.cap = cap;
    o.keys = ctx->alloc<uint64_t>(std::max<uint64_t>(cap, 1));
    o.bufs.push_back(o.keys);
    o.left = alloc_cols(ctx, l, cap, o.bufs);
    o.right = alloc_cols(ctx, r, cap, o.bufs);
    return o;
}

void free_output(phj_ctx *, OutBuffers &o) { o.release(); }

OutSink sink_of(const OutBuffers &o, unsigned long long *cursor) {
    OutSink s{};
    s.cursor = cursor;
    s.cap = o.cap;
    s.keys = o.keys;
    s.left = o.left;
    s.right = o.right;
    return s;
}

template <class T> T read_scalar(phj_ctx *ctx, const T *dev) {
    T v{};
    HIPCHK(hipMemcpyAsync(&v, dev, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    return v;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
    }
    ~EventPair() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    float ms() const {
        float m = 0;
        HIPCHK(hipEventElapsedTime(&m, a, b));
        return m;
    }
};

void fill_result_cols(phj_join_result *res, const OutBuffers &o, uint64_t rows) {
    res->out_rows = rows;
    res->keys = o.keys;
    res->n_left = o.left.n;
    res->n_right = o.right.n;
    for (uint32_t c = 0; c < o.left.n; ++c) {
        res->left_types[c] = o.left.types[c];
        res->left[c] = o.left.ptr[c];
    }
    for (uint32_t c = 0; c < o.right.n; ++c) {
        res->right_types[c] = o.right.types[c];
        res->right[c] = o.right.ptr[c];
    }
}

void set_phase(phj_phase_timing &p, double ms, uint64_t rows) {
    p.wall_ms = ms;
    p.ns_per_row = rows == 0 ? 0.0 : ms * 1e6 / static_cast<double>(rows);
}

uint64_t own_result(phj_ctx *ctx, OutBuffers &o) {
    const uint64_t h = ctx->next_handle++;
    ctx->owned[h] = o.bufs;
    o.bufs.clear();
    return h;
}

// ---------------------------------------------------------------- leaf join
struct LeafJoinPlan {
    uint32_t load_pct = 50;
    uint32_t chunk_rows = 0; // 0: twice the mean probe leaf, at least 16384
    uint32_t threads = 512;
    uint32_t lds_limit_rows = 0;
    bool force_global = false;
    bool two_phase = false; // count pass + exact-offset emit even without repeated build keys
};

LeafJoinPlan plan_of(const phj_tuning &t) {
    LeafJoinPlan p;
    if (t.ht_load_pct)
        p.load_pct = std::min<uint32_t>(std::max<uint32_t>(t.ht_load_pct, 10), 90);
    if (t.probe_chunk_rows)
        p.chunk_rows = t.probe_chunk_rows;
    if (t.join_threads)
        p.threads = t.join_threads;
    p.lds_limit_rows = t.lds_leaf_rows;
    p.force_global = (t.flags & PHJ_TUNE_FORCE_GLOBAL_HT) != 0;
    p.two_phase = (t.flags & PHJ_TUNE_TWO_PHASE_JOIN) != 0;
    return p;
}

// Scatter load/store variant of a tuning record: nontemporal loads unless the
// caller asks for cached ones (the input is read once per pass).
uint32_t radix_flags(const phj_tuning &t) {
    return ((t.flags & PHJ_TUNE_CACHED_LOADS) ? 0u : kRadixNtLoads) | ((t.flags & PHJ_TUNE_NT_STORES) ? kRadixNtStores : 0u) |
           ((t.flags & PHJ_TUNE_HASH_EVERY_PASS) ? kRadixHashEveryPass : 0u) |
           ((t.flags & PHJ_TUNE_ATOMIC_RANK) ? kRadixAtomicRank : 0u);
}

// Largest leaf (build rows) the LDS join paths take for a build side whose
// largest leaf has max_leaf rows: two workgroups per CU when the largest leaf
// allows, otherwise one; 0 when HBM tables are forced.
uint32_t lds_leaf_cap(const LeafJoinPlan &plan, uint32_t max_leaf) {
    uint32_t cap = leaf_join_lds_max_rows(150u * 1024u, plan.load_pct);
    const uint32_t two_per_cu = leaf_join_lds_max_rows(78u * 1024u, plan.load_pct);
    if (max_leaf <= two_per_cu)
        cap = two_per_cu;
    if (plan.lds_limit_rows)
        cap = std::min(cap, plan.lds_limit_rows);
    return plan.force_global ? 0 : std::min(max_leaf, cap);
}

// Fused-join tables of every leaf of a build side as 16-byte cells in HBM,
// built once for joins that probe it many times with few probe rows per leaf
// (PHJ best-effort slices); lookups read them in place.
struct LeafTableImages {
    const unsigned char *img = nullptr;
    uint64_t stride = 0;
    const uint32_t *side = nullptr;
    uint32_t max_leaf = 0;
};

// Joins partitioned R and S leaf by leaf. When every leaf fits LDS the fused
// join runs first: bound() is then |S| and emit() writes the rows in one
// launch, unless a build key repeats, in which case emit() returns kRetry
// after setting up the two-phase plan. The two-phase plan counts the exact
// output rows at construction (LDS leaves by a count pass, oversize leaves
// through HBM tables probed in count-only mode); emit() then writes them from
// a given base row of an output buffer.
// After an HBM table's build (with ht.dups set up before it): when a key
// repeated, every flagged cell's chain is laid out once as a contiguous group
// of the build columns `cols`, so no probe walks next[] (src/JoinOps.h:260-271
// emits the same rows per match).
void group_chains(phj_ctx *ctx, Scratch &scr, GlobalHT &ht, HtMode mode, uint64_t cells, const ColSet &cols) {
    if (!read_scalar(ctx, ht.dups))
        return;
    uint32_t *goff = scr.get<uint32_t>(cells + 1);
    ht_group_count(ctx->stream, ht, mode, cells, goff, ctx->tm());
    uint32_t *gtmp = scr.get<uint32_t>(radix_scan_blocks(cells + 1) + 1);
    scan_u32(ctx->stream, goff, cells + 1, nullptr, 1, gtmp, ctx->tm(), "ht_group");
    const uint32_t grows = read_scalar(ctx, goff + cells);
    OutColSet grp{};
    grp.n = cols.n;
    for (uint32_t c = 0; c < cols.n; ++c) {
        grp.types[c] = cols.types[c];
        grp.ptr[c] = scr.get<unsigned char>(std::max<uint64_t>(grows, 1) * type_size(cols.types[c]));
    }
    ht.grp = as_in(grp);
    ht.grouped = 1;
    uint32_t *rows = scr.get<uint32_t>(grows);
    ht_group_fill(ctx->stream, ht, mode, cells, goff, rows, ctx->tm());
    gather_perm(ctx->stream, cols, grp, rows, grows, ctx->tm());
}

class LeafJoin {
public:
    // `img`: prebuilt fused tables of R's leaves (then R has no repeated key
    // and every leaf fits LDS), or null.
    LeafJoin(phj_ctx *ctx, const Partitioned &R, const Partitioned &S, const LeafJoinPlan &plan,
             const LeafTableImages *img = nullptr)
        : ctx_(ctx), scr_(ctx), plan_(plan), R_(R), S_(S) {
        const uint32_t n_leaves = static_cast<uint32_t>(R.parts);
        uint32_t max_leaf;
        if (img) {
            max_leaf = img->max_leaf;
        } else if (R.max_part >= 0) {
            max_leaf = static_cast<uint32_t>(R.max_part);
        } else {
            uint32_t *dmax = scr_.get<uint32_t>(1);
            leaf_max_rows(ctx->stream, R.off, n_leaves, dmax);
            max_leaf = read_scalar(ctx, dmax);
        }
        lds_max_ = lds_leaf_cap(plan, max_leaf);

        a_.r_keys = R.keys;
        a_.r_cols = as_in(R.cols);
        a_.r_off = R.off;
        a_.s_keys = S.keys;
        a_.s_cols = as_in(S.cols);
        a_.s_off = S.off;
        a_.n_leaves = n_leaves;
        a_.lds_max_rows = std::max<uint32_t>(lds_max_, 1);
        a_.load_pct = plan.load_pct;
        max_leaf_ = max_leaf;
        // Every leaf in LDS: try the fused join first. It is exact when no
        // build key repeats (each probe row then has at most one match, so
        // S.rows bounds the output); a repeated key sends emit() back here.
        fused_ = !plan.two_phase && lds_max_ > 0 && max_leaf <= lds_max_;
        if (fused_ && img) {
            // fixed probe-row ranges against the prebuilt tables: no work plan
            a_.table_img = img->img;
            a_.img_stride = img->stride;
            a_.img_side = img->side;
            while ((1ull << leaf_bits_) < R.parts)
                ++leaf_bits_;
            bound_ = S.rows;
            return;
        }
        if (fused_) {
            plan_work(leaf_join_fused_chunk());
            bound_ = S.rows;
            return;
        }
        plan_two_phase();
    }

    // Sentinel of emit(): the fused attempt met a repeated build key; the
    // two-phase plan is now in place, bound() is exact, call emit() again.
    static constexpr uint64_t kRetry = ~0ull;

    // Output rows emit() may write: exact for the two-phase join, S.rows for
    // the fused one.
    uint64_t bound() const { return bound_; }
    bool fused() const { return fused_; }
    // The probe side is narrow (keys and row indices `rowid` into the
    // unpartitioned columns `src`): matched rows gather their probe columns
    // from there. Fused join only; a repeated build key makes emit() return
    // kRetry without a two-phase plan.
    void use_narrow(const uint32_t *rowid, const ColSet &src) {
        a_.s_rowid = rowid;
        a_.s_src = src;
    }
    double big_build_ms() const { return big_build_ms_; }

    // Writes the join's rows from row `base` of `out` (capacity >= base +
    // bound()) and returns how many.
    uint64_t emit(const OutBuffers &out, uint64_t base) {
        if (fused_) {
            // cursor (word 0) and repeated-key flag (word 1) read back together
            unsigned long long *cf = scr_.get<unsigned long long>(2);
            launch_fill_u64(ctx_->stream, reinterpret_cast<uint64_t *>(cf), 1, base);
            launch_fill_u64(ctx_->stream, reinterpret_cast<uint64_t *>(cf) + 1, 1, 0);
            a_.out = sink_of(out, cf);
            a_.chain_flag = reinterpret_cast<uint32_t *>(cf + 1);
            if (a_.table_img)
                prebuilt_join(ctx_->stream, a_, S_.rows, leaf_bits_, ctx_->tm());
            else
                leaf_join_fused(ctx_->stream, a_, max_work_, ctx_->tm());
            uint64_t h[2];
            HIPCHK(hipMemcpyAsync(h, cf, sizeof(h), hipMemcpyDeviceToHost, ctx_->stream));
            ctx_->sync();
            const uint64_t rows = h[0] - base;
            if (h[1] == 0)
                return rows;
            fused_ = false;
            if (a_.s_rowid)
                return kRetry; // a narrow probe side has no payload to emit from: the caller re-partitions
            plan_two_phase();
            return kRetry;
        }
        if (lds_max_ > 0 && lds_rows_ > 0) {
            a_.out = sink_of(out, nullptr);
            a_.out_base = base;
            leaf_emit(ctx_->stream, a_, max_work_, chains_, ctx_->tm());
        }
        if (nb_ > 0 && big_rows_ > 0) {
            unsigned long long *cursor = scr_.get<unsigned long long>(1);
            launch_fill_u64(ctx_->stream, reinterpret_cast<uint64_t *>(cursor), 1, base + lds_rows_);
            ht_probe_emit(ctx_->stream, ht_, kHtLeaves, hp_, probe_big_, as_in(R_.cols), as_in(S_.cols),
                          sink_of(out, cursor), ctx_->tm(), "leaf_ht_probe");
        }
        return lds_rows_ + big_rows_;
    }

private:
    // Work items of the LDS leaves: probe chunks of `chunk` rows.
    void plan_work(uint32_t chunk) {
        const uint32_t n_leaves = a_.n_leaves;
        a_.chunk_rows = chunk;
        uint32_t *work_first = scr_.get<uint32_t>(n_leaves + 1);
        uint32_t *scan_tmp = scr_.get<uint32_t>(radix_scan_blocks(n_leaves + 1) + 1);
        max_work_ = leaf_max_work(S_.rows, n_leaves, chunk);
        uint32_t *item_leaf = scr_.get<uint32_t>(max_work_);
        a_.work_first = work_first;
        a_.item_leaf = item_leaf;
        leaf_work_plan(ctx_->stream, a_, work_first, scan_tmp, item_leaf, ctx_->tm());
    }

    // Count pass (LDS leaves: probe kernel and scan; oversize leaves: HBM
    // tables probed in count-only mode) giving exact output offsets.
    void plan_two_phase() {
        const uint32_t n_leaves = a_.n_leaves;
        if (lds_max_ > 0) {
            plan_work(plan_.chunk_rows ? plan_.chunk_rows
                                       : static_cast<uint32_t>(std::max<uint64_t>(
                                             16384, 2 * (S_.rows / std::max<uint32_t>(n_leaves, 1)))));
            uint32_t *counts = scr_.get<uint32_t>(max_work_ + 1);
            launch_fill_u32(ctx_->stream, counts, max_work_ + 1, 0);
            a_.item_count = counts;
            a_.heads = scr_.get<uint16_t>(S_.rows);
            uint32_t *chain_first = scr_.get<uint32_t>(n_leaves + 1);
            uint32_t *ctmp = scr_.get<uint32_t>(radix_scan_blocks(n_leaves + 1) + 1);
            leaf_chain_slots(ctx_->stream, a_, chain_first, ctmp);
            a_.chain_first = chain_first;
            // per work item: its row groups and their lengths, leaf rows each
            a_.next16 = scr_.get<uint16_t>(2 * static_cast<uint64_t>(std::max<uint32_t>(read_scalar(ctx_, chain_first + n_leaves), 1)));
            a_.chain_flag = scr_.get<uint32_t>(1);
            launch_fill_u32(ctx_->stream, a_.chain_flag, 1, 0);
            leaf_probe(ctx_->stream, a_, max_work_, plan_.threads, ctx_->tm());
            uint32_t *tmp = scr_.get<uint32_t>(radix_scan_blocks(max_work_ + 1) + 1);
            scan_u32(ctx_->stream, counts, max_work_ + 1, nullptr, 1, tmp, ctx_->tm(), "join_plan");
            a_.item_off = counts;
            lds_rows_ = read_scalar(ctx_, counts + max_work_);
            chains_ = read_scalar(ctx_, a_.chain_flag) != 0;
        }
        if (max_leaf_ > lds_max_)
            plan_big_leaves(n_leaves);
        bound_ = lds_rows_ + big_rows_;
    }

    // HBM tables for the oversize leaves: table b = b-th listed leaf.
    void plan_big_leaves(uint32_t n_leaves) {
        uint32_t *leaf_of = scr_.get<uint32_t>(n_leaves);
        uint32_t *cnt = scr_.get<uint32_t>(1);
        leaf_big_list(ctx_->stream, R_.off, S_.off, n_leaves, lds_max_, leaf_of, cnt);
        nb_ = read_scalar(ctx_, cnt);
        if (nb_ == 0)
            return;
        std::vector<uint32_t> h_leaf(nb_), h_roff(n_leaves + 1), h_soff(n_leaves + 1);
        HIPCHK(hipMemcpyAsync(h_leaf.data(), leaf_of, nb_ * 4ull, hipMemcpyDeviceToHost, ctx_->stream));
        HIPCHK(hipMemcpyAsync(h_roff.data(), R_.off, (n_leaves + 1) * 4ull, hipMemcpyDeviceToHost, ctx_->stream));
        HIPCHK(hipMemcpyAsync(h_soff.data(), S_.off, (n_leaves + 1) * 4ull, hipMemcpyDeviceToHost, ctx_->stream));
        ctx_->sync();
        std::sort(h_leaf.begin(), h_leaf.end());
        std::vector<uint64_t> cell_off(nb_ + 1, 0);
        std::vector<uint32_t> rfirst(nb_ + 1, 0), sfirst(nb_ + 1, 0);
        for (uint32_t b = 0; b < nb_; ++b) {
            const uint32_t L = h_leaf[b];
            const uint64_t nr = h_roff[L + 1] - h_roff[L];
            uint64_t cap = 2;
            while (cap < nr * 100 / plan_.load_pct + 1)
                cap <<= 1;
            cell_off[b + 1] = cell_off[b] + cap;
            rfirst[b + 1] = rfirst[b] + static_cast<uint32_t>(nr);
            sfirst[b + 1] = sfirst[b] + (h_soff[L + 1] - h_soff[L]);
        }
        uint64_t *d_cell_off = scr_.get<uint64_t>(nb_ + 1);
        uint32_t *d_leaf = scr_.get<uint32_t>(nb_);
        uint32_t *d_rfirst = scr_.get<uint32_t>(nb_ + 1);
        uint32_t *d_sfirst = scr_.get<uint32_t>(nb_ + 1);
        HIPCHK(hipMemcpyAsync(d_cell_off, cell_off.data(), (nb_ + 1) * 8ull, hipMemcpyHostToDevice, ctx_->stream));
        HIPCHK(hipMemcpyAsync(d_leaf, h_leaf.data(), nb_ * 4ull, hipMemcpyHostToDevice, ctx_->stream));
        HIPCHK(hipMemcpyAsync(d_rfirst, rfirst.data(), (nb_ + 1) * 4ull, hipMemcpyHostToDevice, ctx_->stream));
        HIPCHK(hipMemcpyAsync(d_sfirst, sfirst.data(), (nb_ + 1) * 4ull, hipMemcpyHostToDevice, ctx_->stream));
        ht_.cells = scr_.get<Cell16>(cell_off[nb_]);
        ht_.cell_off = d_cell_off;
        ht_.next = scr_.get<uint32_t>(R_.rows);
        ht_.side_head = scr_.get<uint32_t>(nb_);
        ht_.n_tables = nb_;
        ht_.block_rows = 1;
        ht_.build_seg_off = R_.off;
        ht_.dups = scr_.get<uint32_t>(1);
        launch_fill_u32(ctx_->stream, ht_.dups, 1, 0);
        EventPair ev;
        HIPCHK(hipEventRecord(ev.a, ctx_->stream));
        ht_init(ctx_->stream, ht_, cell_off[nb_], R_.rows, ctx_->tm());
        HtRows hr{};
        hr.keys = R_.keys;
        hr.rows = R_.rows;
        hr.seg_off = R_.off;
        hr.leaf_of = d_leaf;
        hr.first = d_rfirst;
        ht_build(ctx_->stream, ht_, kHtLeaves, hr, rfirst[nb_], ctx_->tm(), "leaf_ht_build");
        group_chains(ctx_, scr_, ht_, kHtLeaves, cell_off[nb_], as_in(R_.cols));
        HIPCHK(hipEventRecord(ev.b, ctx_->stream));
        hp_.keys = S_.keys;
        hp_.rows = S_.rows;
        hp_.seg_off = S_.off;
        hp_.leaf_of = d_leaf;
        hp_.first = d_sfirst;
        probe_big_ = sfirst[nb_];
        // count-only probe: a sink of capacity 0 counts without writing
        unsigned long long *cursor = scr_.get<unsigned long long>(1);
        launch_fill_u64(ctx_->stream, reinterpret_cast<uint64_t *>(cursor), 1, 0);
        OutSink counter{};
        counter.cursor = cursor;
        counter.cap = 0;
        ht_probe_emit(ctx_->stream, ht_, kHtLeaves, hp_, probe_big_, as_in(R_.cols), as_in(S_.cols), counter, ctx_->tm(),
                      "leaf_ht_count");
        big_rows_ = read_scalar(ctx_, reinterpret_cast<const uint64_t *>(cursor));
        big_build_ms_ = ev.ms();
    }

    phj_ctx *ctx_;
    Scratch scr_;
    LeafJoinPlan plan_;
    const Partitioned &R_, &S_;
    LeafJoinArgs a_{};
    uint32_t lds_max_ = 0;
    uint64_t max_work_ = 0;
    uint64_t lds_rows_ = 0;
    bool chains_ = false;
    uint32_t nb_ = 0;
    GlobalHT ht_{};
    HtRows hp_{};
    uint64_t probe_big_ = 0;
    uint64_t big_rows_ = 0;
    double big_build_ms_ = 0;
    bool fused_ = false;
    uint64_t bound_ = 0;
    uint32_t max_leaf_ = 0;
    uint32_t leaf_bits_ = 0; // prebuilt tables: log2 of the leaf count
};

} // namespace

namespace {

void validate_rel(const phj_relation *r) {
    if (r == nullptr)
        invalid("null relation");
    if (r->n_cols > PHJ_MAX_COLS)
        invalid("at most 8 payload columns");
    for (uint32_t c = 0; c < r->n_cols; ++c)
        tsize(r->cols[c].type);
}

// Re-allocates the output with `cap` rows, keeping the first `used` rows.
void grow_output(phj_ctx *ctx, OutBuffers &out, uint64_t used, uint64_t cap) {
    OutBuffers bigger = alloc_output(ctx, as_in(out.left), as_in(out.right), cap);
    launch_copy(ctx->stream, bigger.keys, out.keys, used * 8, nullptr);
    for (uint32_t c = 0; c < out.left.n; ++c)
        launch_copy(ctx->stream, bigger.left.ptr[c], out.left.ptr[c], used * type_size(out.left.types[c]), nullptr);
    for (uint32_t c = 0; c < out.right.n; ++c)
        launch_copy(ctx->stream, bigger.right.ptr[c], out.right.ptr[c], used * type_size(out.right.types[c]), nullptr);
    free_output(ctx, out);
    out = std::move(bigger);
}

struct ResultBuilder {
    phj_ctx *ctx;
    phj_join_result *res;
    uint64_t build_rows, probe_rows;
    void finish(OutBuffers &out, uint64_t rows) {
        fill_result_cols(res, out, rows);
        res->peak_mem_bytes = ctx->arena.peak();
        res->handle = own_result(ctx, out);
    }
};

// Joins with the output capacity grown to the exact row count on overflow.
template <class F> uint64_t join_with_capacity(phj_ctx *ctx, const ColSet &l, const ColSet &r, uint64_t cap, OutBuffers &out, F &&run) {
    Scratch scr(ctx);
    unsigned long long *cursor = scr.get<unsigned long long>(1);
    out = alloc_output(ctx, l, r, cap);
    launch_fill_u64(ctx->stream, reinterpret_cast<uint64_t *>(cursor), 1, 0);
    uint64_t n = run(out, cursor);
    if (n > out.cap) {
        free_output(ctx, out);
        out = alloc_output(ctx, l, r, n);
        launch_fill_u64(ctx->stream, reinterpret_cast<uint64_t *>(cursor), 1, 0);
        n = run(out, cursor);
    }
    return n;
}

// ---- PHJ-PURE: src/PHJ.cpp:19-150 ----
// Probe half of PHJ-PURE: partitions S like the build side (same passes)
// and joins leaf by leaf against the partitioned build side RP. `e2e` was
// started by the caller; `ebs` brackets the build shuffle (null: resident).
void phj_pure_probe(phj_ctx *ctx, const Partitioned &RP, const ColSet &r_cols, const DevRel &S,
                    const std::vector<uint32_t> &passes, const phj_tuning &tun, EventPair &e2e, EventPair *ebs,
                    Held &keep, phj_join_result *res) {
    const LeafJoinPlan plan = plan_of(tun);
    EventPair eps, ej;
    Partitioned R = RP;
    Scratch scr(ctx);
    if (R.max_part < 0)
        max_part_start(ctx, R, scr);
    double big_build_ms = 0;
    OutBuffers out;
    uint64_t n = 0;
    // PHJ_TUNE_NARROW_PROBE: the probe side is partitioned as (key, row index)
    // only, 12 bytes a row instead of its full width, and the fused join
    // gathers the probe columns of matched rows from the unpartitioned
    // relation. It needs every leaf in LDS and no repeated build key; failing
    // either, the join runs again with the full-width probe side.
    bool narrow = (tun.flags & PHJ_TUNE_NARROW_PROBE) != 0 && S.rows > 0;
    for (bool done = false; !done;) {
        DevRel Sp = S;
        Held sk(ctx);
        if (narrow) {
            Sp.cols = ColSet{};
            Sp.cols.n = 1;
            Sp.cols.types[0] = kTypeRowIndex;
        }
        HIPCHK(hipEventRecord(eps.a, ctx->stream));
        const Partitioned SP = radix_partition(ctx, Sp, nullptr, 1, 0, passes, sk, "probe", radix_flags(tun));
        HIPCHK(hipEventRecord(eps.b, ctx->stream));
        HIPCHK(hipEventRecord(ej.a, ctx->stream));
        if (R.max_part < 0)
            max_part_wait(ctx, R); // the probe passes are queued behind it
        LeafJoin lj(ctx, R, SP, plan);
        if (narrow) {
            if (!lj.fused()) {
                narrow = false;
                continue;
            }
            lj.use_narrow(static_cast<const uint32_t *>(SP.cols.ptr[0]), S.cols);
        }
        out = alloc_output(ctx, r_cols, S.cols, lj.bound());
        while ((n = lj.emit(out, 0)) == LeafJoin::kRetry && !narrow) {
            free_output(ctx, out);
            out = alloc_output(ctx, r_cols, S.cols, lj.bound());
        }
        if (n == LeafJoin::kRetry) { // narrow join met a repeated build key
            free_output(ctx, out);
            narrow = false;
            continue;
        }
        big_build_ms = lj.big_build_ms();
        for (void *b : sk)
            keep.push_back(b);
        sk.clear();
        done = true;
    }
    HIPCHK(hipEventRecord(ej.b, ctx->stream));
    HIPCHK(hipEventRecord(e2e.b, ctx->stream));
    ctx->sync();
    keep.free_all();
    set_phase(res->build_shuffle, ebs ? ebs->ms() : 0.0, RP.rows);
    set_phase(res->probe_shuffle, eps.ms(), S.rows);
    // Leaf build and probe run fused in one kernel; the fused time is
    // reported as probe, and build carries only separately built HBM tables.
    set_phase(res->build, big_build_ms, RP.rows);
    set_phase(res->probe, ej.ms() - big_build_ms, S.rows);
    res->e2e_wall_ms = e2e.ms();
    ResultBuilder{ctx, res, RP.rows, S.rows}.finish(out, n);
}

// Leaf bits of a PHJ-PURE join. The output does not depend on the leaves, so
// a configuration whose leaves would average more than kJoinLeafRows build
// rows (too large for an LDS table) is refined to that size, up to the device
// maximum; {4,4} on 2^27 build rows then joins like {8,8}. An explicit LDS
// leaf limit or forced HBM tables keep the configuration as given.
constexpr uint64_t kJoinLeafRows = 2048;
// Joins run at most 2^21 leaves: enough for ~2,048-row leaves at the 2^32-row
// relation limit, and every leaf-join grid (work items = leaves + probe
// chunks, up to 1,024 threads each) stays below HIP's 2^32 work-items per
// launch dimension.
constexpr uint32_t kDeviceMaxTotalBits = 21;

// Configurations of 25..32 total bits, which the reference accepts
// (src/CLI.cpp:350-355, src/RadixPartition.cpp:556-557), join through the
// device maximum of 24 bits: the output is the same for any leaf split.
uint32_t join_total_bits(uint32_t requested, uint64_t build_rows, const phj_tuning &tun) {
    requested = std::min(requested, kDeviceMaxTotalBits);
    if (tun.lds_leaf_rows != 0 || (tun.flags & PHJ_TUNE_FORCE_GLOBAL_HT) != 0)
        return requested;
    uint32_t b = requested;
    while (b < kDeviceMaxTotalBits && (build_rows >> b) > kJoinLeafRows)
        ++b;
    return b;
}

void run_phj_pure(phj_ctx *ctx, const phj_relation &b, const phj_relation &p, const uint8_t *pb, uint32_t np,
                  const phj_tuning &tun, phj_join_result *res) {
    validate_rel(&b);
    validate_rel(&p);
    const uint32_t total = join_total_bits(validate_pass_bits(pb, np), b.rows, tun);
    const std::vector<uint32_t> passes = device_passes(total);
    std::memset(res, 0, sizeof(*res));
    Scratch scr(ctx);
    const DevRel R = stage_relation(ctx, scr, b);
    const DevRel S = stage_relation(ctx, scr, p);
    ctx->arena.reset_peak();

    EventPair e2e, ebs;
    HIPCHK(hipEventRecord(e2e.a, ctx->stream));
    HIPCHK(hipEventRecord(ebs.a, ctx->stream));
    Held keep(ctx);
    const Partitioned RP = radix_partition(ctx, R, nullptr, 1, 0, passes, keep, "build", radix_flags(tun));
    HIPCHK(hipEventRecord(ebs.b, ctx->stream));
    phj_pure_probe(ctx, RP, R.cols, S, passes, tun, e2e, &ebs, keep, res);
}

// ---- CHJ: src/CHJ.cpp:16-119 with 256 lock-free sub-tables ----
void run_chj(phj_ctx *ctx, const phj_relation &b, const phj_relation &p, phj_join_result *res) {
    validate_rel(&b);
    validate_rel(&p);
    std::memset(res, 0, sizeof(*res));
    Scratch scr(ctx);
    const DevRel R = stage_relation(ctx, scr, b);
    const DevRel S = stage_relation(ctx, scr, p);
    ctx->arena.reset_peak();
    constexpr uint32_t kTables = 256;
    EventPair e2e, eb, ep;
    HIPCHK(hipEventRecord(e2e.a, ctx->stream));
    HIPCHK(hipEventRecord(eb.a, ctx->stream));
    uint32_t *counts = scr.get<uint32_t>(kTables);
    ht_chj_plan(ctx->stream, R.keys, R.rows, kTables, counts, ctx->tm());
    std::vector<uint32_t> hc(kTables);
    HIPCHK(hipMemcpyAsync(hc.data(), counts, kTables * 4, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    std::vector<uint64_t> cell_off(kTables + 1, 0);
    for (uint32_t t = 0; t < kTables; ++t) {
        uint64_t cap = 2; // JoinHashTable::reserve: bit_ceil(2 * rows), src/HashTable.h:72-77
        while (cap < 2ull * hc[t])
            cap <<= 1;
        cell_off[t + 1] = cell_off[t] + cap;
    }
    uint64_t *d_off = scr.get<uint64_t>(kTables + 1);
    HIPCHK(hipMemcpyAsync(d_off, cell_off.data(), (kTables + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
    GlobalHT ht{};
    // Cells of 32 or 64 bytes carry a one-row key's build columns inline, so
    // its probes read them from the cell they just fetched instead of from
    // random build rows; wider payloads keep 16-byte cells.
    ht.cell_shift = ht_wide_shift(R.cols);
    if (ht.cell_shift)
        ht.pay = R.cols;
    ht.cells = scr.get<Cell16>(cell_off[kTables] << (ht.cell_shift ? ht.cell_shift - 4 : 0));
    ht.cell_off = d_off;
    ht.next = scr.get<uint32_t>(R.rows);
    ht.side_head = scr.get<uint32_t>(kTables);
    ht.n_tables = kTables;
    ht.block_rows = PHJ_PIPELINE_BLOCK_ROWS;
    ht.dups = scr.get<uint32_t>(1);
    launch_fill_u32(ctx->stream, ht.dups, 1, 0);
    ht_init(ctx->stream, ht, cell_off[kTables], R.rows, ctx->tm());
    HtRows hr{};
    hr.keys = R.keys;
    hr.rows = R.rows;
    ht_build(ctx->stream, ht, kHtChj, hr, R.rows, ctx->tm(), "chj_build");
    group_chains(ctx, scr, ht, kHtChj, cell_off[kTables], R.cols);
    HIPCHK(hipEventRecord(eb.b, ctx->stream));
    HIPCHK(hipEventRecord(ep.a, ctx->stream));
    HtRows hp{};
    hp.keys = S.keys;
    hp.rows = S.rows;
    // Without repeated keys |S| bounds the output; with them a count-only
    // probe (no chain walks: the groups give each key's rows) sizes it.
    uint64_t cap = std::max<uint64_t>(S.rows, 1);
    if (ht.grouped) {
        unsigned long long *cnt = scr.get<unsigned long long>(1);
        launch_fill_u64(ctx->stream, reinterpret_cast<uint64_t *>(cnt), 1, 0);
        OutSink counter{};
        counter.cursor = cnt;
        counter.cap = 0;
        ht_probe_emit(ctx->stream, ht, kHtChj, hp, S.rows, R.cols, S.cols, counter, ctx->tm(), "chj_count");
        cap = std::max<uint64_t>(read_scalar(ctx, reinterpret_cast<const uint64_t *>(cnt)), 1);
    }
    OutBuffers out;
    const uint64_t n = join_with_capacity(ctx, R.cols, S.cols, cap, out,
                                          [&](const OutBuffers &o, unsigned long long *cur) {
                                              ht_probe_emit(ctx->stream, ht, kHtChj, hp, S.rows, R.cols, S.cols,
                                                            sink_of(o, cur), ctx->tm(), "chj_probe");
                                              return read_scalar(ctx, reinterpret_cast<const uint64_t *>(cur));
                                          });
    HIPCHK(hipEventRecord(ep.b, ctx->stream));
    HIPCHK(hipEventRecord(e2e.b, ctx->stream));
    ctx->sync();
    set_phase(res->build, eb.ms(), R.rows);
    set_phase(res->probe, ep.ms(), S.rows);
    res->e2e_wall_ms = e2e.ms();
    ResultBuilder{ctx, res, R.rows, S.rows}.finish(out, n);
}

} // namespace

// A build side partitioned once and kept in HBM (device-resident BlockStore
// equivalent): probes against it skip the build shuffle.
struct phj_resident {
    phj_ctx *owner = nullptr;
    // Probes running against this build side (calls on any context); release
    // is refused while it is nonzero.
    std::atomic<uint32_t> inflight{0};
    Partitioned RP;
    ColSet r_cols{}; // column types (pointers unused)
    std::vector<uint32_t> passes;
    std::vector<void *> bufs;
    uint64_t bytes = 0;
};

namespace {

void register_build(phj_ctx *ctx, const phj_relation &b, const uint8_t *pb, uint32_t np, const phj_tuning &tun,
                    phj_resident *r) {
    validate_rel(&b);
    const uint32_t total = join_total_bits(validate_pass_bits(pb, np), b.rows, tun);
    r->owner = ctx;
    r->passes = device_passes(total);
    Scratch scr(ctx);
    const DevRel R = stage_relation(ctx, scr, b);
    r->RP = radix_partition(ctx, R, nullptr, 1, 0, r->passes, r->bufs, "build", radix_flags(tun));
    r->r_cols = R.cols;
    uint64_t w = 8;
    for (uint32_t c = 0; c < R.cols.n; ++c)
        w += type_size(R.cols.types[c]);
    r->bytes = R.rows * w + (r->RP.parts + 1) * 4;
    max_part_start(ctx, r->RP, scr); // known once, so probes from any context skip reading it back
    max_part_wait(ctx, r->RP);
    ctx->sync();
}

void probe_resident(phj_ctx *ctx, const phj_resident &r, const phj_relation &p, const phj_tuning &tun,
                    phj_join_result *res) {
    validate_rel(&p);
    std::memset(res, 0, sizeof(*res));
    Scratch scr(ctx);
    const DevRel S = stage_relation(ctx, scr, p);
    ctx->arena.reset_peak();
    EventPair e2e;
    HIPCHK(hipEventRecord(e2e.a, ctx->stream));
    Held keep(ctx);
    phj_pure_probe(ctx, r.RP, r.r_cols, S, r.passes, tun, e2e, nullptr, keep, res);
}

// Builds the tables of every leaf of RP into `img` (buffers go to `keep`) for
// probes of up to `probe_rows` rows each. False, with nothing kept, when those
// probes hold a quarter of the largest leaf or more per leaf (each work item
// then builds its table in LDS: one sequential read of the leaf's keys beats
// a random HBM read per probe row), when a leaf exceeds the LDS cap, when a
// key repeats, or when the plan asks for another join path. The build time
// is added to `build_ms`.
bool build_table_images(phj_ctx *ctx, const Partitioned &RP, const LeafJoinPlan &plan, uint64_t probe_rows,
                        std::vector<void *> &keep, LeafTableImages &img, double &build_ms) {
    if (plan.two_phase || plan.force_global || RP.parts == 0)
        return false;
    const uint32_t n_leaves = static_cast<uint32_t>(RP.parts);
    Scratch scr(ctx);
    uint32_t *word = scr.get<uint32_t>(2); // [0] largest leaf, [1] repeated-key flag
    leaf_max_rows(ctx->stream, RP.off, n_leaves, word);
    const uint32_t max_leaf = read_scalar(ctx, word);
    if (max_leaf == 0 || lds_leaf_cap(plan, max_leaf) < max_leaf ||
        4 * probe_rows >= static_cast<uint64_t>(n_leaves) * max_leaf)
        return false;
    EventPair ev;
    HIPCHK(hipEventRecord(ev.a, ctx->stream));
    const uint64_t stride = leaf_table_bytes(max_leaf, plan.load_pct);
    unsigned char *tables = ctx->alloc<unsigned char>(stride * n_leaves);
    uint32_t *side = ctx->alloc<uint32_t>(n_leaves);
    LeafJoinArgs a{};
    a.r_keys = RP.keys;
    a.r_off = RP.off;
    a.n_leaves = n_leaves;
    a.lds_max_rows = max_leaf;
    a.load_pct = plan.load_pct;
    launch_fill_u32(ctx->stream, word + 1, 1, 0);
    leaf_table_images(ctx->stream, a, tables, stride, side, word + 1, ctx->tm());
    HIPCHK(hipEventRecord(ev.b, ctx->stream));
    if (read_scalar(ctx, word + 1) != 0) {
        ctx->arena.free(tables);
        ctx->arena.free(side);
        return false;
    }
    build_ms += ev.ms();
    keep.push_back(tables);
    keep.push_back(side);
    img.img = tables;
    img.stride = stride;
    img.side = side;
    img.max_leaf = max_leaf;
    return true;
}

// ---- PHJ (best-effort partitioning): src/PHJBep.cpp:113-790 ----
// The reference's schedule, run by one device scheduler (the host side of the
// context's stream) instead of `threads` CPU workers:
//   - the build side is shuffled to full leaf depth and every leaf's table is
//     built up front in HBM when the build side allows (src/PHJBep.cpp:157-231):
//     single-row cells without repeated keys, group cells with them (the
//     key's rows adjacent), on leaves refined past the configuration's when
//     those would not fit an LDS build;
//   - the probe side is scattered by its first pass bits, one batch at a time,
//     into pass-1 buckets ("unrefined", src/PHJBep.cpp:574-606). A batch is a
//     quarter of the unrefined high-water mark, the device's unit of work where
//     the reference scatters one 10,000-row block; the next batch's scatter is
//     queued before the host waits for this one's bucket counts;
//   - after each batch, when unrefined rows reach the high-water mark, the
//     largest buckets are refined through the remaining passes into leaf
//     buffers until the low-water mark holds (src/PHJBep.cpp:508-533, :397-444);
//     all buckets chosen at one trigger go through one segmented pass sequence;
//   - when buffered leaf rows reach their high-water mark, leaves are evicted
//     round-robin from a cursor (empty leaves skipped) until their low-water
//     mark holds: each evicted leaf's buffered probe rows are joined against
//     its table and dropped (src/PHJBep.cpp:453-505, :535-568); all leaves of
//     one trigger are joined in one launch sequence, which against tables
//     built up front reserves its output on a device cursor and never waits;
//   - at the end every residual bucket is refined and every leaf drained
//     (src/PHJBep.cpp:633-695).
// Marks (src/PHJBep.cpp:186-190, one worker): unrefined high water budget/2,
// low water a half of it; leaf high water 5/8 of the budget, low water 5/8 of
// that. Rows are buffered in "runs" (one batch's pass-1 output, one refine's
// leaf output) and read back by gathering their still-live row ranges, so no
// buffer is ever grown in place. The device keeps at most kBepMaxRuns runs of
// each kind: beyond that the buckets (leaves) still held by the oldest run
// are refined (evicted) before the next batch (refine). Counters follow the
// reference: refinements = buckets refined, evictions = leaves evicted
// (drain included), peak = buffered leaf rows x row bytes at its largest
// after a batch, eviction overhead = choosing victims and gathering their
// rows. bep_build_skip_retries counts failed leaf try-locks between CPU
// workers; one scheduler takes no locks, so it is always 0.
constexpr size_t kBepMaxRuns = 8;

struct BepRun {
    Held bufs;
    DevRel rel;                 // the run's rows: keys + columns
    std::vector<uint32_t> off;  // host copy of its group offsets (groups + 1)
    std::vector<uint8_t> done;  // per group: rows already refined/evicted
    std::vector<int32_t> qidx;  // leaf runs: bucket -> its index in the refined set, -1 if absent
    uint64_t live = 0;          // rows not yet consumed
    explicit BepRun(phj_ctx *c) : bufs(c) {}
    uint32_t len(uint32_t g) const { return off[g + 1] - off[g]; }
};

// Leaf tables of the build side for the eviction rounds: built once in HBM
// (single-row cells without repeated keys, group cells with them), or none
// (each round's work items build theirs in LDS).
struct BepTables {
    enum Kind { kLds, kUnique, kGroup } kind = kLds;
    LeafTableImages unique;
    GroupImage group{};
    uint32_t leaf_bits = 0;
};

// Group images of a partitioned build side: cells, side groups and the
// grouped build columns, in `keep`. False when a leaf is too large for the
// LDS build or the plan asks for LDS or HBM leaf tables.
bool build_group_images(phj_ctx *ctx, const Partitioned &RP, const LeafJoinPlan &plan, uint32_t leaf_bits,
                        uint32_t max_leaf, Held &keep, GroupImage &g, double &build_ms) {
    constexpr size_t kLdsLimit = 150u * 1024u;
    if (plan.two_phase || plan.force_global || RP.parts == 0 || max_leaf == 0 ||
        group_image_lds_bytes(max_leaf, plan.load_pct) > kLdsLimit ||
        (plan.lds_limit_rows && max_leaf > plan.lds_limit_rows))
        return false;
    const uint32_t n_leaves = static_cast<uint32_t>(RP.parts);
    EventPair ev;
    HIPCHK(hipEventRecord(ev.a, ctx->stream));
    const uint64_t stride = leaf_table_bytes(max_leaf, plan.load_pct) / sizeof(Cell16);
    Cell16 *cells = ctx->alloc<Cell16>(stride * n_leaves);
    keep.push_back(cells);
    uint32_t *side = ctx->alloc<uint32_t>(2ull * n_leaves);
    keep.push_back(side);
    OutColSet gcols = alloc_cols(ctx, as_in(RP.cols), RP.rows, keep);
    {
        Scratch scr(ctx);
        uint32_t *perm = scr.get<uint32_t>(RP.rows);
        group_images(ctx->stream, RP.keys, RP.off, n_leaves, max_leaf, plan.load_pct, cells, stride, side, perm,
                     ctx->tm());
        gather_perm(ctx->stream, as_in(RP.cols), gcols, perm, RP.rows, ctx->tm());
    }
    HIPCHK(hipEventRecord(ev.b, ctx->stream));
    g.cells = cells;
    g.stride = stride;
    g.side = side;
    g.r_off = RP.off;
    g.cols = as_in(gcols);
    g.leaf_bits = leaf_bits;
    g.load_pct = plan.load_pct;
    HIPCHK(hipEventSynchronize(ev.b));
    build_ms += ev.ms();
    return true;
}

class BepSchedule {
public:
    // RP: the build side partitioned for the tables (the schedule's leaves
    // for the LDS join; possibly finer for tables built up front).
    BepSchedule(phj_ctx *ctx, const Partitioned &RP, const DevRel &S, uint32_t p1_bits, const std::vector<uint32_t> &rest,
                uint64_t budget, const LeafJoinPlan &plan, uint32_t flags, const BepTables &tables, OutBuffers &out)
        : ctx_(ctx), RP_(RP), S_(S), p1_(p1_bits), rest_(rest), plan_(plan), flags_(flags), tab_(tables), out_(out),
          keep_(ctx) {
        P_ = 1u << p1_;
        uint32_t total = p1_;
        for (uint32_t b : rest_)
            total += b;
        leaves_ = 1u << total;
        lpp_ = leaves_ / P_;
        rb_ = 8;
        for (uint32_t c = 0; c < S.cols.n; ++c)
            rb_ += type_size(S.cols.types[c]);
        const uint64_t tile = radix_tile_rows();
        u_hi_ = std::max<uint64_t>(budget / 2 / rb_, tile);
        u_lo_ = u_hi_ / 2;
        l_hi_ = std::max<uint64_t>(budget * 5 / 8 / rb_, 1);
        l_lo_ = l_hi_ * 5 / 8;
        batch_ = std::max<uint64_t>((u_hi_ / 4 + tile - 1) / tile * tile, tile);
        bucket_rows_.assign(P_, 0);
        leaf_rows_.assign(leaves_, 0);
        for (auto &b : scat_off_)
            b = ctx_->pinned_io.get<uint32_t>(P_ + 1);
        leaf_off_ = ctx_->pinned_io.get<uint32_t>(static_cast<uint64_t>(leaves_) + 1);
        cursor_host_ = ctx_->pinned_io.get<unsigned long long>(1);
        if (tab_.kind != BepTables::kLds) {
            dcursor_ = ctx_->alloc<unsigned long long>(1);
            keep_.push_back(dcursor_);
            launch_fill_u64(ctx_->stream, reinterpret_cast<uint64_t *>(dcursor_), 1, 0);
        }
    }

    void run() {
        const uint64_t nb = (S_.rows + batch_ - 1) / batch_;
        // the next batch's scatter is queued before the host waits for this
        // batch's bucket counts, so the device does not idle meanwhile
        std::unique_ptr<Pending> cur = nb ? launch_scatter(0) : nullptr;
        for (uint64_t b = 0; b < nb; ++b) {
            std::unique_ptr<Pending> next = b + 1 < nb ? launch_scatter(b + 1) : nullptr;
            while (batch_runs_.size() >= kBepMaxRuns)
                refine(held_by(*batch_runs_.front(), P_, [](uint32_t g) { return g; }));
            finish_scatter(*cur);
            cur = std::move(next);
            peak_ = std::max(peak_, global_leaf_ * rb_);
            // src/PHJBep.cpp:508-533: drain own unrefined to the low-water mark, argmax first
            if (unrefined_ >= u_hi_) {
                std::vector<uint32_t> order(P_);
                for (uint32_t q = 0; q < P_; ++q)
                    order[q] = q;
                std::stable_sort(order.begin(), order.end(),
                                 [&](uint32_t x, uint32_t y) { return bucket_rows_[x] > bucket_rows_[y]; });
                std::vector<uint32_t> Q;
                uint64_t left = unrefined_;
                for (uint32_t q : order) {
                    if (left < u_lo_ || bucket_rows_[q] == 0)
                        break;
                    Q.push_back(q);
                    left -= bucket_rows_[q];
                }
                refine(Q);
            }
            evict_to_low_water();
        }
        // src/PHJBep.cpp:678-695: refine every residual bucket, then drain every leaf
        std::vector<uint32_t> Q;
        uint64_t rows = 0;
        for (uint32_t q = 0; q < P_; ++q) {
            if (bucket_rows_[q] == 0)
                continue;
            if (!Q.empty() && rows + bucket_rows_[q] > u_hi_) {
                refine(Q);
                evict_to_low_water();
                Q.clear();
                rows = 0;
            }
            Q.push_back(q);
            rows += bucket_rows_[q];
        }
        if (!Q.empty())
            refine(Q);
        std::vector<uint32_t> E;
        rows = 0;
        for (uint32_t L = 0; L < leaves_; ++L) {
            if (leaf_rows_[L] == 0)
                continue;
            if (!E.empty() && rows + leaf_rows_[L] > std::max(l_hi_, u_hi_)) {
                evict(E);
                E.clear();
                rows = 0;
            }
            E.push_back(L);
            rows += leaf_rows_[L];
        }
        if (!E.empty())
            evict(E);
        if (dcursor_) {
            HIPCHK(hipMemcpyAsync(cursor_host_, dcursor_, 8, hipMemcpyDeviceToHost, ctx_->stream));
            ctx_->sync();
            n_out_ = *cursor_host_;
            if (n_out_ > out_.cap)
                throw std::runtime_error("best-effort join wrote past its output bound");
        } else {
            ctx_->sync();
        }
        upload_.reset();
        fold_events();
    }

    uint64_t out_rows() const { return n_out_; }
    uint64_t evictions() const { return evictions_; }
    uint64_t refinements() const { return refinements_; }
    uint64_t peak_bytes() const { return peak_; }
    double shuffle_ms() const { return shuffle_ms_; }
    double probe_ms() const { return probe_ms_; }
    double overhead_ms() const { return overhead_ms_; }
    double big_build_ms() const { return big_build_ms_; }

private:
    // A batch scatter in flight: its run and the pinned copy of its offsets.
    struct Pending {
        std::unique_ptr<BepRun> run;
        uint32_t *off;
        hipEvent_t done = nullptr;
        ~Pending() {
            if (done)
                (void)hipEventDestroy(done);
        }
    };

    EventPair &timed(std::vector<std::unique_ptr<EventPair>> &v) {
        if (free_ev_.empty()) {
            v.push_back(std::make_unique<EventPair>());
        } else {
            v.push_back(std::move(free_ev_.back()));
            free_ev_.pop_back();
        }
        HIPCHK(hipEventRecord(v.back()->a, ctx_->stream));
        return *v.back();
    }

    // After a wait that covers everything queued so far: the phase times of
    // the recorded event pairs are added up and the pairs reused.
    void fold_events() {
        auto fold = [&](std::vector<std::unique_ptr<EventPair>> &v, double &acc) {
            for (auto &e : v) {
                acc += e->ms();
                free_ev_.push_back(std::move(e));
            }
            v.clear();
        };
        fold(shuffle_ev_, shuffle_ms_);
        fold(probe_ev_, probe_ms_);
        fold(gather_ev_, overhead_ms_);
    }

    // Groups of `run` still holding rows, mapped to bucket or leaf ids.
    template <class F> std::vector<uint32_t> held_by(const BepRun &run, uint32_t groups, F id) const {
        std::vector<uint32_t> v;
        for (uint32_t g = 0; g < groups; ++g)
            if (!run.done[g] && run.len(g) > 0)
                v.push_back(id(g));
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        return v;
    }

    // Pass-1 scatter of probe rows of batch b into a new batch run; its
    // bucket offsets are copied back into pinned memory.
    std::unique_ptr<Pending> launch_scatter(uint64_t b) {
        const uint64_t a = b * batch_, n = std::min(batch_, S_.rows - a);
        auto p = std::make_unique<Pending>();
        p->run = std::make_unique<BepRun>(ctx_);
        p->off = scat_off_[b % 2];
        EventPair &ev = timed(shuffle_ev_);
        DevRel slice = S_;
        slice.rows = n;
        slice.keys = S_.keys + a;
        for (uint32_t c = 0; c < S_.cols.n; ++c)
            slice.cols.ptr[c] = static_cast<const uint8_t *>(S_.cols.ptr[c]) + a * type_size(S_.cols.types[c]);
        const Partitioned P1 = radix_partition(ctx_, slice, nullptr, 1, 0, {p1_}, p->run->bufs, "probe", flags_);
        HIPCHK(hipEventRecord(ev.b, ctx_->stream));
        p->run->rel.rows = n;
        p->run->rel.keys = P1.keys;
        p->run->rel.cols = as_in(P1.cols);
        HIPCHK(hipMemcpyAsync(p->off, P1.off, (P_ + 1) * 4ull, hipMemcpyDeviceToHost, ctx_->stream));
        HIPCHK(hipEventCreateWithFlags(&p->done, hipEventDisableTiming));
        HIPCHK(hipEventRecord(p->done, ctx_->stream));
        return p;
    }

    void finish_scatter(Pending &p) {
        HIPCHK(hipEventSynchronize(p.done));
        BepRun &run = *p.run;
        run.off.assign(p.off, p.off + P_ + 1);
        run.done.assign(P_, 0);
        run.live = run.rel.rows;
        for (uint32_t q = 0; q < P_; ++q)
            bucket_rows_[q] += run.len(q);
        unrefined_ += run.rel.rows;
        batch_runs_.push_back(std::move(p.run));
    }

    // Rows of group g of each run in `runs` (for which pick() gives g, or -1),
    // gathered into one relation: groups in the caller's order, each group's
    // rows from the oldest run first. Returns the staged relation (buffers in
    // `keep`) and the row count per caller item. The segment list goes up
    // through pinned memory; nothing waits.
    template <class Pick>
    DevRel gather(std::vector<std::unique_ptr<BepRun>> &runs, uint32_t items, Pick pick, Held &keep,
                  std::vector<uint64_t> &item_rows) {
        std::vector<uint64_t> start{0}, row;
        std::vector<uint32_t> src;
        item_rows.assign(items, 0);
        for (uint32_t i = 0; i < items; ++i)
            for (uint32_t k = 0; k < runs.size(); ++k) {
                const int64_t g = pick(*runs[k], i);
                if (g < 0 || runs[k]->done[g] || runs[k]->len(static_cast<uint32_t>(g)) == 0)
                    continue;
                const uint32_t len = runs[k]->len(static_cast<uint32_t>(g));
                src.push_back(k);
                row.push_back(runs[k]->off[g]);
                start.push_back(start.back() + len);
                item_rows[i] += len;
            }
        return stage_segments(runs, start, row, src, keep);
    }

    // Rows [start[j], start[j + 1]) of the staged relation are rows row[j] ..
    // of run src[j]. One pinned upload of the lists, one gather launch.
    DevRel stage_segments(std::vector<std::unique_ptr<BepRun>> &runs, const std::vector<uint64_t> &seg_start,
                          const std::vector<uint64_t> &seg_row, const std::vector<uint32_t> &seg_src, Held &keep) {
        const uint64_t n_seg = seg_src.size();
        // one upload: start (n_seg + 1) | row (n_seg) as u64, src (n_seg) as u32, then the sources
        const uint64_t bytes_u64 = (2 * n_seg + 1) * 8, bytes_src = (n_seg * 4 + 15) & ~uint64_t{15};
        const uint64_t bytes = bytes_u64 + bytes_src + runs.size() * sizeof(GatherSource);
        unsigned char *h = static_cast<unsigned char *>(upload_.bytes(bytes));
        uint64_t *start = reinterpret_cast<uint64_t *>(h);
        uint64_t *row = start + n_seg + 1;
        uint32_t *src = reinterpret_cast<uint32_t *>(h + bytes_u64);
        GatherSource *srcs = reinterpret_cast<GatherSource *>(h + bytes_u64 + bytes_src);
        std::memcpy(start, seg_start.data(), (n_seg + 1) * 8);
        std::memcpy(row, seg_row.data(), n_seg * 8);
        std::memcpy(src, seg_src.data(), n_seg * 4);
        for (size_t k = 0; k < runs.size(); ++k) {
            srcs[k] = GatherSource{};
            srcs[k].keys = runs[k]->rel.keys;
            for (uint32_t c = 0; c < S_.cols.n; ++c)
                srcs[k].cols[c] = runs[k]->rel.cols.ptr[c];
        }
        DevRel d;
        d.rows = start[n_seg];
        d.cols.n = S_.cols.n;
        uint64_t *keys = ctx_->alloc<uint64_t>(std::max<uint64_t>(d.rows, 1));
        keep.push_back(keys);
        OutColSet cols = alloc_cols(ctx_, S_.cols, d.rows, keep);
        d.keys = keys;
        d.cols = as_in(cols);
        if (d.rows == 0)
            return d;
        Scratch scr(ctx_); // freed blocks are reused in stream order only
        unsigned char *dev = scr.get<unsigned char>(bytes);
        HIPCHK(hipMemcpyAsync(dev, h, bytes, hipMemcpyHostToDevice, ctx_->stream));
        gather_segments(ctx_->stream, reinterpret_cast<const uint64_t *>(dev), static_cast<uint32_t>(n_seg),
                        reinterpret_cast<const uint32_t *>(dev + bytes_u64),
                        reinterpret_cast<const uint64_t *>(dev) + n_seg + 1,
                        reinterpret_cast<const GatherSource *>(dev + bytes_u64 + bytes_src),
                        static_cast<uint32_t>(runs.size()), d.rows, keys, cols, ctx_->tm());
        return d;
    }

    // src/PHJBep.cpp:397-444 refinePartition, for every bucket of Q at once:
    // the buckets' rows from every batch run, through the remaining passes
    // (refine_to_leaves mode), into one new leaf run. The host waits for the
    // new run's leaf offsets (the eviction decisions need them); everything
    // queued before is then done, so the upload memory is recycled.
    void refine(const std::vector<uint32_t> &Q) {
        if (Q.empty())
            return;
        while (leaf_runs_.size() >= kBepMaxRuns) {
            const BepRun &old = *leaf_runs_.front();
            std::vector<uint32_t> E;
            for (uint32_t q = 0; q < P_; ++q)
                if (old.qidx[q] >= 0)
                    for (uint32_t d = 0; d < lpp_; ++d) {
                        const uint32_t g = static_cast<uint32_t>(old.qidx[q]) * lpp_ + d;
                        if (!old.done[g] && old.len(g) > 0)
                            E.push_back(q * lpp_ + d);
                    }
            evict(E);
        }
        EventPair &ev = timed(shuffle_ev_);
        auto leaf = std::make_unique<BepRun>(ctx_);
        Held staging(ctx_);
        std::vector<uint64_t> q_rows;
        const DevRel st = gather(
            batch_runs_, static_cast<uint32_t>(Q.size()), [&](const BepRun &, uint32_t i) { return int64_t(Q[i]); },
            staging, q_rows);
        const uint32_t nq = static_cast<uint32_t>(Q.size());
        const size_t n_off = rest_.empty() ? nq + 1 : static_cast<size_t>(nq) * lpp_ + 1;
        if (rest_.empty()) {
            // one pass in all: the buckets are the leaves (src/RadixPartition.cpp:437-447)
            leaf->rel = st;
            for (void *p : staging)
                leaf->bufs.push_back(p);
            staging.clear();
            leaf->off.assign(nq + 1, 0);
            for (uint32_t i = 0; i < nq; ++i)
                leaf->off[i + 1] = leaf->off[i] + static_cast<uint32_t>(q_rows[i]);
            HIPCHK(hipEventRecord(ev.b, ctx_->stream));
            ctx_->sync();
        } else {
            uint32_t *seg = upload_.get<uint32_t>(nq + 1);
            seg[0] = 0;
            for (uint32_t i = 0; i < nq; ++i)
                seg[i + 1] = seg[i] + static_cast<uint32_t>(q_rows[i]);
            Scratch scr(ctx_);
            uint32_t *dseg = scr.get<uint32_t>(nq + 1);
            HIPCHK(hipMemcpyAsync(dseg, seg, (nq + 1) * 4ull, hipMemcpyHostToDevice, ctx_->stream));
            const Partitioned LP = radix_partition(ctx_, st, dseg, nq, p1_, rest_, leaf->bufs, "refine", flags_);
            staging.free_all();
            leaf->rel.rows = st.rows;
            leaf->rel.keys = LP.keys;
            leaf->rel.cols = as_in(LP.cols);
            HIPCHK(hipEventRecord(ev.b, ctx_->stream));
            HIPCHK(hipMemcpyAsync(leaf_off_, LP.off, n_off * 4, hipMemcpyDeviceToHost, ctx_->stream));
            ctx_->sync();
            leaf->off.assign(leaf_off_, leaf_off_ + n_off);
        }
        upload_.reset();
        fold_events();
        leaf->done.assign(static_cast<size_t>(nq) * lpp_, 0);
        leaf->qidx.assign(P_, -1);
        for (uint32_t i = 0; i < nq; ++i) {
            const uint32_t q = Q[i];
            leaf->qidx[q] = static_cast<int32_t>(i);
            for (uint32_t d = 0; d < lpp_; ++d) {
                const uint32_t cnt = leaf->len(i * lpp_ + d);
                leaf_rows_[q * lpp_ + d] += cnt;
                global_leaf_ += cnt;
            }
            // the bucket's rows have left every batch run
            for (auto &r : batch_runs_)
                if (!r->done[q]) {
                    r->done[q] = 1;
                    r->live -= r->len(q);
                }
            unrefined_ -= bucket_rows_[q];
            bucket_rows_[q] = 0;
        }
        leaf->live = st.rows;
        refinements_ += nq;
        batch_runs_.erase(std::remove_if(batch_runs_.begin(), batch_runs_.end(),
                                         [](const std::unique_ptr<BepRun> &r) { return r->live == 0; }),
                          batch_runs_.end());
        leaf_runs_.push_back(std::move(leaf));
    }

    // src/PHJBep.cpp:535-568: evict round-robin until below the low-water mark.
    void evict_to_low_water() {
        if (global_leaf_ < l_hi_)
            return;
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<uint32_t> E;
        uint64_t left = global_leaf_;
        for (uint32_t scanned = 0; scanned < leaves_ && left >= l_lo_; ++scanned) {
            const uint32_t L = cursor_;
            cursor_ = cursor_ + 1 == leaves_ ? 0 : cursor_ + 1;
            if (leaf_rows_[L] == 0)
                continue; // the reference's relaxed pre-check skips empty leaves
            E.push_back(L);
            left -= leaf_rows_[L];
        }
        // cursor order is ascending but for one wrap past the last leaf
        for (size_t i = 1; i < E.size(); ++i)
            if (E[i] < E[i - 1]) {
                std::rotate(E.begin(), E.begin() + static_cast<std::ptrdiff_t>(i), E.end());
                break;
            }
        overhead_ms_ += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        evict(E);
    }

    // src/PHJBep.cpp:453-505 tryEvictOneLeaf, for every leaf of E at once: the
    // leaves' buffered rows from every leaf run, joined against their tables,
    // then dropped. Against tables built up front (single-row or group
    // images) the join reserves its rows on a device cursor and nothing
    // waits; otherwise the leaves' work items build LDS tables (LeafJoin).
    void evict(const std::vector<uint32_t> &E) {
        if (E.empty())
            return;
        Held staging(ctx_);
        EventPair &gev = timed(gather_ev_);
        if (tab_.kind == BepTables::kLds) {
            std::vector<uint64_t> l_rows;
            const DevRel st = gather(
                leaf_runs_, static_cast<uint32_t>(E.size()),
                [&](const BepRun &r, uint32_t i) {
                    const int32_t qi = r.qidx[E[i] / lpp_];
                    return qi < 0 ? int64_t(-1) : int64_t(qi) * lpp_ + E[i] % lpp_;
                },
                staging, l_rows);
            HIPCHK(hipEventRecord(gev.b, ctx_->stream));
            evict_lds(E, st, l_rows);
            consume(E);
        } else {
            const DevRel st = gather_evicted(E, staging);
            HIPCHK(hipEventRecord(gev.b, ctx_->stream));
            join_images(st);
        }
        for (const uint32_t L : E) {
            global_leaf_ -= leaf_rows_[L];
            leaf_rows_[L] = 0;
        }
        evictions_ += E.size();
        leaf_runs_.erase(std::remove_if(leaf_runs_.begin(), leaf_runs_.end(),
                                        [](const std::unique_ptr<BepRun> &r) { return r->live == 0; }),
                         leaf_runs_.end());
    }

    // Marks the groups of the leaves of E consumed in every leaf run.
    void consume(const std::vector<uint32_t> &E) {
        for (const uint32_t L : E)
            for (auto &r : leaf_runs_) {
                const int32_t qi = r->qidx[L / lpp_];
                if (qi < 0)
                    continue;
                const uint32_t g = static_cast<uint32_t>(qi) * lpp_ + L % lpp_;
                if (!r->done[g]) {
                    r->done[g] = 1;
                    r->live -= r->len(g);
                }
            }
    }

    // The buffered rows of the leaves of E (ascending) from every leaf run,
    // consumed as they are listed. The join against tables built up front
    // takes rows in any order, so each run's groups of one interval of
    // consecutive leaves (adjacent in the run) merge into few segments.
    DevRel gather_evicted(const std::vector<uint32_t> &E, Held &keep) {
        std::vector<uint64_t> start{0}, row;
        std::vector<uint32_t> src;
        for (size_t i = 0; i < E.size();) {
            size_t j = i;
            while (j + 1 < E.size() && E[j + 1] == E[j] + 1)
                ++j;
            const uint32_t a = E[i], b = E[j];
            for (uint32_t k = 0; k < leaf_runs_.size(); ++k) {
                BepRun &r = *leaf_runs_[k];
                for (uint32_t q = a / lpp_; q <= b / lpp_; ++q) {
                    const int32_t qi = r.qidx[q];
                    if (qi < 0)
                        continue;
                    const uint32_t d0 = q == a / lpp_ ? a % lpp_ : 0, d1 = q == b / lpp_ ? b % lpp_ : lpp_ - 1;
                    for (uint32_t d = d0; d <= d1; ++d) {
                        const uint32_t g = static_cast<uint32_t>(qi) * lpp_ + d;
                        if (r.done[g])
                            continue;
                        r.done[g] = 1;
                        const uint32_t len = r.len(g);
                        if (len == 0)
                            continue;
                        r.live -= len;
                        const size_t n = src.size();
                        if (n > 0 && src[n - 1] == k && row[n - 1] + (start[n] - start[n - 1]) == r.off[g]) {
                            start[n] += len;
                        } else {
                            src.push_back(k);
                            row.push_back(r.off[g]);
                            start.push_back(start[n] + len);
                        }
                    }
                }
            }
            i = j + 1;
        }
        return stage_segments(leaf_runs_, start, row, src, keep);
    }

    // Rows of the leaves of E, staged by gather_evicted, joined against the
    // tables built up front; the join reserves its rows on the device cursor.
    void join_images(const DevRel &st) {
        if (tab_.kind == BepTables::kUnique) {
            EventPair &ev = timed(probe_ev_);
            LeafJoinArgs a{};
            a.r_keys = RP_.keys;
            a.r_cols = as_in(RP_.cols);
            a.r_off = RP_.off;
            a.s_keys = st.keys;
            a.s_cols = st.cols;
            a.n_leaves = static_cast<uint32_t>(RP_.parts);
            a.load_pct = plan_.load_pct;
            a.table_img = tab_.unique.img;
            a.img_stride = tab_.unique.stride;
            a.img_side = tab_.unique.side;
            a.out = sink_of(out_, dcursor_);
            prebuilt_join(ctx_->stream, a, st.rows, tab_.leaf_bits, ctx_->tm());
            HIPCHK(hipEventRecord(ev.b, ctx_->stream));
        } else {
            EventPair &ev = timed(probe_ev_);
            group_join(ctx_->stream, tab_.group, st.keys, st.cols, st.rows, sink_of(out_, dcursor_), ctx_->tm());
            HIPCHK(hipEventRecord(ev.b, ctx_->stream));
        }
    }

    // The staged rows as a partitioned relation over all leaves, joined leaf
    // by leaf with LDS tables built by each work item.
    void evict_lds(const std::vector<uint32_t> &E, const DevRel &st, const std::vector<uint64_t> &l_rows) {
        std::vector<uint32_t> off(static_cast<size_t>(leaves_) + 1, 0);
        for (size_t i = 0, L = 0; L < leaves_; ++L) {
            off[L + 1] = off[L];
            if (i < E.size() && E[i] == L)
                off[L + 1] += static_cast<uint32_t>(l_rows[i++]);
        }
        Scratch scr(ctx_);
        uint32_t *doff = scr.get<uint32_t>(off.size());
        HIPCHK(hipMemcpyAsync(doff, off.data(), off.size() * 4, hipMemcpyHostToDevice, ctx_->stream));
        Partitioned SP;
        SP.rows = st.rows;
        SP.keys = const_cast<uint64_t *>(st.keys);
        SP.cols.n = st.cols.n;
        for (uint32_t c = 0; c < st.cols.n; ++c) {
            SP.cols.types[c] = st.cols.types[c];
            SP.cols.ptr[c] = const_cast<void *>(st.cols.ptr[c]);
        }
        SP.off = doff;
        SP.parts = leaves_;
        EventPair ev;
        HIPCHK(hipEventRecord(ev.a, ctx_->stream));
        {
            LeafJoinPlan plan = plan_;
            plan.two_phase = plan.two_phase || repeated_keys_;
            LeafJoin lj(ctx_, RP_, SP, plan, nullptr);
            uint64_t add;
            while (true) {
                if (n_out_ + lj.bound() > out_.cap)
                    grow_output(ctx_, out_, n_out_, std::max(n_out_ + lj.bound(), 2 * out_.cap));
                add = lj.emit(out_, n_out_);
                if (add != LeafJoin::kRetry)
                    break;
                repeated_keys_ = true; // later rounds go straight to the two-phase join
            }
            n_out_ += add;
            big_build_ms_ += lj.big_build_ms();
            probe_ms_ -= lj.big_build_ms();
        }
        HIPCHK(hipEventRecord(ev.b, ctx_->stream));
        ctx_->sync();
        probe_ms_ += ev.ms();
    }

    phj_ctx *ctx_;
    const Partitioned &RP_;
    const DevRel &S_;
    uint32_t p1_;
    std::vector<uint32_t> rest_;
    LeafJoinPlan plan_;
    uint32_t flags_;
    const BepTables &tab_;
    OutBuffers &out_;
    Held keep_;
    PinnedArena upload_;
    uint32_t *scat_off_[2] = {nullptr, nullptr};
    uint32_t *leaf_off_ = nullptr;
    unsigned long long *cursor_host_ = nullptr;
    unsigned long long *dcursor_ = nullptr;
    uint32_t P_ = 0, leaves_ = 0, lpp_ = 0;
    uint64_t rb_ = 0, u_hi_ = 0, u_lo_ = 0, l_hi_ = 0, l_lo_ = 0, batch_ = 0;
    std::vector<uint64_t> bucket_rows_, leaf_rows_;
    uint64_t unrefined_ = 0, global_leaf_ = 0;
    uint32_t cursor_ = 0;
    std::vector<std::unique_ptr<BepRun>> batch_runs_, leaf_runs_;
    std::vector<std::unique_ptr<EventPair>> shuffle_ev_, probe_ev_, gather_ev_, free_ev_;
    bool repeated_keys_ = false;
    uint64_t n_out_ = 0, evictions_ = 0, refinements_ = 0, peak_ = 0;
    double shuffle_ms_ = 0, probe_ms_ = 0, overhead_ms_ = 0, big_build_ms_ = 0;
};

void run_phj_bep(phj_ctx *ctx, const phj_relation &b, const phj_relation &p, const uint8_t *pb, uint32_t np,
                 uint64_t budget_bytes, const phj_tuning &tun, phj_join_result *res) {
    validate_rel(&b);
    validate_rel(&p);
    const uint32_t total = std::min(validate_pass_bits(pb, np), kDeviceMaxTotalBits);
    // pass 1 keeps the configuration's first pass bits (at most one device pass)
    const uint32_t p1 = std::min({static_cast<uint32_t>(pb[0]), kDeviceMaxPassBits, total});
    const std::vector<uint32_t> rest = total > p1 ? device_passes(total - p1) : std::vector<uint32_t>{};
    std::vector<uint32_t> passes{p1};
    passes.insert(passes.end(), rest.begin(), rest.end());
    std::memset(res, 0, sizeof(*res));
    Scratch scr(ctx);
    const DevRel R = stage_relation(ctx, scr, b);
    const DevRel S = stage_relation(ctx, scr, p);
    ctx->arena.reset_peak();
    const LeafJoinPlan plan = plan_of(tun);
    uint64_t row_bytes = 8;
    for (uint32_t c = 0; c < S.cols.n; ++c)
        row_bytes += type_size(S.cols.types[c]);

    EventPair e2e, ebs;
    HIPCHK(hipEventRecord(e2e.a, ctx->stream));
    HIPCHK(hipEventRecord(ebs.a, ctx->stream));
    Held keep(ctx);
    // The tables' leaves: the configuration's, or finer when its leaves would
    // average more than an LDS leaf's rows (as PHJ-PURE refines them). The
    // schedule's buckets and leaves, and so its counters, stay the
    // configuration's; a probe row finds its table from the top bits of its
    // hash, which the finer leaves refine.
    const uint32_t join_bits = std::max(total, join_total_bits(total, b.rows, tun));
    Held fine(ctx);
    Partitioned RP = radix_partition(ctx, R, nullptr, 1, 0, join_bits > total ? device_passes(join_bits) : passes,
                                     join_bits > total ? fine : keep, "build", radix_flags(tun));
    HIPCHK(hipEventRecord(ebs.b, ctx->stream));
    // Eager leaf tables (src/PHJBep.cpp:157-231), built once in HBM: single-
    // row cells when every leaf fits LDS, no key repeats and an eviction round
    // holds few probe rows per leaf; group cells when keys repeat. Otherwise
    // each round's work items build theirs in LDS (on the configuration's
    // leaves).
    double build_ms = 0;
    BepTables tables;
    tables.leaf_bits = join_bits;
    const uint64_t round_rows = std::max<uint64_t>(budget_bytes * 5 / 8 / row_bytes * 3 / 8, 1);
    uint64_t out_cap = std::max<uint64_t>(S.rows, 1);
    if (build_table_images(ctx, RP, plan, round_rows, keep, tables.unique, build_ms)) {
        tables.kind = BepTables::kUnique;
    } else if (RP.rows > 0 && S.rows > 0) {
        uint32_t max_leaf = static_cast<uint32_t>(std::max<int64_t>(RP.max_part, 0));
        if (RP.max_part < 0) {
            Scratch s2(ctx);
            uint32_t *dmax = s2.get<uint32_t>(1);
            leaf_max_rows(ctx->stream, RP.off, static_cast<uint32_t>(RP.parts), dmax);
            max_leaf = read_scalar(ctx, dmax);
        }
        if (build_group_images(ctx, RP, plan, join_bits, max_leaf, keep, tables.group, build_ms)) {
            tables.kind = BepTables::kGroup;
            // the output's exact size: every probe row's group length, summed
            Scratch s2(ctx);
            unsigned long long *cnt = s2.get<unsigned long long>(1);
            launch_fill_u64(ctx->stream, reinterpret_cast<uint64_t *>(cnt), 1, 0);
            group_count(ctx->stream, tables.group, S.keys, S.rows, cnt, ctx->tm());
            out_cap = std::max<uint64_t>(read_scalar(ctx, cnt), 1);
        }
    }
    if (join_bits > total) {
        if (tables.kind == BepTables::kLds) {
            // no tables up front: the rounds join leaf by leaf on the configuration's leaves
            fine.free_all();
            HIPCHK(hipEventRecord(ebs.a, ctx->stream));
            RP = radix_partition(ctx, R, nullptr, 1, 0, passes, keep, "build", radix_flags(tun));
            HIPCHK(hipEventRecord(ebs.b, ctx->stream));
        } else {
            keep.insert(keep.end(), fine.begin(), fine.end());
            fine.clear();
        }
    }
    OutBuffers out = alloc_output(ctx, R.cols, S.cols, out_cap);
    uint64_t n_out = 0;
    {
        BepSchedule sched(ctx, RP, S, p1, rest, budget_bytes, plan, radix_flags(tun), tables, out);
        sched.run();
        HIPCHK(hipEventRecord(e2e.b, ctx->stream));
        ctx->sync();
        set_phase(res->build, build_ms + sched.big_build_ms(), R.rows);
        set_phase(res->probe_shuffle, sched.shuffle_ms(), S.rows);
        set_phase(res->probe, sched.probe_ms(), S.rows);
        set_phase(res->eviction_overhead, sched.overhead_ms(), S.rows);
        res->bep_evictions = sched.evictions();
        res->bep_refinements = sched.refinements();
        res->bep_peak_bytes = sched.peak_bytes();
        n_out = sched.out_rows();
    }
    ctx->pinned_io.reset();
    keep.free_all();
    set_phase(res->build_shuffle, ebs.ms(), R.rows);
    res->e2e_wall_ms = e2e.ms();
    res->bep_budget_mib = budget_bytes >> 20;
    res->bep_build_skip_retries = 0;
    ResultBuilder{ctx, res, R.rows, S.rows}.finish(out, n_out);
}

// ---- Batched PHJ-PURE joins ----
// A batch is cut into groups of consecutive joins on device relations with one
// schema, up to kBatchGroupBytes of input. A group whose relations lie back
// to back is one segmented relation per side as it is; otherwise (joins of
// at most kBatchStageBytes each) its relations are first gathered back to
// back. Each join picks its own leaf bits for leaves of about its size
// class's leaf rows (its output does not depend on them; the table below is
// set by phj_batch_autotune); a group runs as many device passes as its
// largest join needs, every pass splitting each join's segments by that
// join's bits, then one fused leaf join over all leaves of all joins writes
// each join's rows to its own region of a shared output (bounded by its probe
// rows, as no build key repeats). A join whose leaves repeat a key or outgrow
// LDS is rerun alone on the plain path after the batch.
//
// Groups are pipelined on the context's one stream: group g+1's partition
// passes are queued before the host waits for group g's leaf sizes (the one
// value the leaf-join launch needs, read back into pinned memory), so the
// device is not left idle while the host plans. Partition buffers go back to
// the pool as soon as their join is queued (stream order keeps them alive
// for it); every upload and readback goes through pinned memory.
constexpr uint64_t kBatchGroupBytes = 8ull << 30;
constexpr uint64_t kBatchStageBytes = 64ull << 20;

// Size class of a join: floor(log2(build rows)) clamped to the table.
uint32_t size_class(uint64_t build_rows) {
    uint32_t c = 0;
    while (c + 1 < kBatchClasses && (build_rows >> (c + 1)) != 0)
        ++c;
    return c;
}

uint32_t batch_bits(uint64_t build_rows, const uint32_t *leaf_rows) {
    const uint64_t target = leaf_rows[size_class(build_rows)];
    uint32_t b = 0;
    while (b < 16 && (build_rows >> b) > target)
        ++b;
    return b;
}

uint64_t rel_bytes(const phj_relation &r) {
    uint64_t w = 8;
    for (uint32_t c = 0; c < r.n_cols; ++c)
        w += type_size(r.cols[c].type);
    return r.rows * w;
}

bool same_schema(const phj_relation &a, const phj_relation &b) {
    if (a.n_cols != b.n_cols || a.on_device != b.on_device)
        return false;
    for (uint32_t c = 0; c < a.n_cols; ++c)
        if (a.cols[c].type != b.cols[c].type)
            return false;
    return true;
}

// b's rows directly follow a's in every column.
bool follows(const phj_relation &a, const phj_relation &b) {
    if (b.keys != a.keys + a.rows)
        return false;
    for (uint32_t c = 0; c < a.n_cols; ++c) {
        const uint64_t w = type_size(a.cols[c].type);
        if (static_cast<const uint8_t *>(b.cols[c].data) != static_cast<const uint8_t *>(a.cols[c].data) + a.rows * w)
            return false;
    }
    return true;
}

// Radix partition of a segmented relation (segment j = join j) where join j
// splits by bits[p][j] in pass p (0 keeps the segment whole). Returns the
// leaves; leaf_join[l] is the join owning leaf l (a join's leaves are
// consecutive). Bit windows are uploaded from pinned memory.
Partitioned radix_partition_joins(phj_ctx *ctx, const DevRel &in, const uint32_t *seg_off, uint32_t n_joins,
                                  const std::vector<std::vector<uint32_t>> &bits, std::vector<void *> &keep,
                                  uint32_t flags, std::vector<uint32_t> &leaf_join) {
    Scratch scr(ctx);
    const uint64_t rows = in.rows;
    std::vector<uint32_t> join_of(n_joins);
    for (uint32_t j = 0; j < n_joins; ++j)
        join_of[j] = j;
    std::vector<uint32_t> consumed(n_joins, 0);
    const uint64_t *cur_keys = in.keys;
    ColSet cur_cols = in.cols;
    const uint32_t *cur_seg = seg_off;
    uint32_t cur_nseg = n_joins;
    Held prev_bufs(ctx);
    Partitioned out;
    out.rows = rows;
    for (size_t p = 0; p < bits.size(); ++p) {
        uint16_t *sb = ctx->pinned_io.get<uint16_t>(cur_nseg);
        uint32_t *first = ctx->pinned_io.get<uint32_t>(static_cast<uint64_t>(cur_nseg) + 1);
        std::vector<uint32_t> next_join;
        uint32_t maxb = 0;
        uint64_t run = 0;
        for (uint32_t sg = 0; sg < cur_nseg; ++sg) {
            const uint32_t j = join_of[sg];
            const uint32_t b = bits[p][j];
            sb[sg] = static_cast<uint16_t>((b ? 64 - consumed[j] - b : 0u) | (b << 8));
            first[sg] = static_cast<uint32_t>(run);
            run += 1ull << b;
            maxb = std::max(maxb, b);
            next_join.insert(next_join.end(), static_cast<size_t>(1) << b, j);
        }
        if (run > 0xFFFFFFF0ull)
            throw PhjError(PHJ_ERR_UNSUPPORTED, "too many partitions in one batch group");
        first[cur_nseg] = static_cast<uint32_t>(run);
        for (uint32_t j = 0; j < n_joins; ++j)
            consumed[j] += bits[p][j];
        uint16_t *dsb = scr.get<uint16_t>(cur_nseg);
        uint32_t *dfirst = scr.get<uint32_t>(static_cast<uint64_t>(cur_nseg) + 1);
        HIPCHK(hipMemcpyAsync(dsb, sb, cur_nseg * 2ull, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(dfirst, first, (cur_nseg + 1ull) * 4, hipMemcpyHostToDevice, ctx->stream));
        const uint64_t B = 1ull << maxb;
        Held bufs(ctx);
        uint64_t *nkeys = ctx->alloc<uint64_t>(std::max<uint64_t>(rows, 1));
        bufs.push_back(nkeys);
        OutColSet ncols = alloc_cols(ctx, cur_cols, rows, bufs);
        uint32_t *nseg = ctx->alloc<uint32_t>(run + 1);
        bufs.push_back(nseg);
        RadixPassScratch rs;
        rs.seg_tile_first = scr.get<uint32_t>(2 * (static_cast<uint64_t>(cur_nseg) + 1));
        rs.hist = scr.get<uint32_t>(radix_max_tiles(rows, cur_nseg) * B);
        rs.block_sums = scr.get<uint32_t>(radix_max_chunks(rows, cur_nseg) * B);
        radix_pass(ctx->stream, cur_keys, cur_cols, nkeys, ncols, cur_seg, cur_nseg, rows, PassBits{dsb, dfirst, 0, maxb},
                   nseg, rs, flags, ctx->tm());
        prev_bufs.free_all();
        prev_bufs.swap(bufs);
        cur_keys = nkeys;
        cur_cols = as_in(ncols);
        cur_seg = nseg;
        cur_nseg = static_cast<uint32_t>(run);
        join_of.swap(next_join);
        out.keys = nkeys;
        out.cols = ncols;
        out.off = nseg;
        out.parts = run;
    }
    for (void *b : prev_bufs)
        keep.push_back(b);
    prev_bufs.clear();
    leaf_join.swap(join_of);
    return out;
}

// Largest leaf (build rows) of each join of a group.
__global__ void k_join_max_leaf(const uint32_t *__restrict__ off, uint32_t n_leaves, const uint32_t *__restrict__ leaf_join,
                                uint32_t *__restrict__ join_max) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n_leaves)
        atomicMax(&join_max[leaf_join[l]], off[l + 1] - off[l]);
}

// One group of a batch in flight.
struct BatchGroup {
    const phj_join_desc *d = nullptr;
    phj_join_result *res = nullptr;
    uint32_t m = 0;
    Held keep;                 // staged inputs and partitions, until the leaf join is queued
    DevRel R, S;
    std::vector<uint64_t> base, end; // output rows of join j: [base[j], base[j] + written)
    uint32_t n_leaves = 0;
    LeafJoinArgs a{};
    uint32_t *join_max_h = nullptr; // pinned readbacks
    unsigned long long *cur_h = nullptr;
    uint32_t *flag_h = nullptr;
    uint32_t lds_max = 0;
    OutBuffers out;
    std::vector<uint8_t> rerun;
    EventPair e2e, ebs, eps, ej;
    explicit BatchGroup(phj_ctx *c) : keep(c) {}
};

// Relation j of a group (build or probe side) as a gather source.
GatherSource source_of(const phj_relation &r) {
    GatherSource g{};
    g.keys = r.keys;
    for (uint32_t c = 0; c < r.n_cols; ++c)
        g.cols[c] = r.cols[c].data;
    return g;
}

// The group's side `which` (0 build, 1 probe) as one relation: in place when
// the joins' relations are back to back, else gathered back to back.
DevRel group_side(phj_ctx *ctx, BatchGroup &g, int which, bool in_place, uint64_t rows) {
    const phj_relation &r0 = which ? g.d[0].probe : g.d[0].build;
    DevRel v;
    v.rows = rows;
    v.cols.n = r0.n_cols;
    for (uint32_t c = 0; c < r0.n_cols; ++c)
        v.cols.types[c] = r0.cols[c].type;
    if (in_place) {
        v.keys = r0.keys;
        for (uint32_t c = 0; c < r0.n_cols; ++c)
            v.cols.ptr[c] = r0.cols[c].data;
        return v;
    }
    uint64_t *seg_start = ctx->pinned_io.get<uint64_t>(g.m + 1);
    uint64_t *seg_row = ctx->pinned_io.get<uint64_t>(g.m);
    uint32_t *seg_src = ctx->pinned_io.get<uint32_t>(g.m);
    GatherSource *srcs = ctx->pinned_io.get<GatherSource>(g.m);
    seg_start[0] = 0;
    uint32_t n_seg = 0; // non-empty relations only
    for (uint32_t j = 0; j < g.m; ++j) {
        const phj_relation &r = which ? g.d[j].probe : g.d[j].build;
        srcs[j] = source_of(r);
        if (r.rows == 0)
            continue;
        seg_start[n_seg + 1] = seg_start[n_seg] + r.rows;
        seg_row[n_seg] = 0;
        seg_src[n_seg] = j;
        ++n_seg;
    }
    Scratch scr(ctx);
    uint64_t *dstart = scr.get<uint64_t>(g.m + 1);
    uint64_t *drow = scr.get<uint64_t>(g.m);
    uint32_t *dsrc = scr.get<uint32_t>(g.m);
    GatherSource *dsrcs = scr.get<GatherSource>(g.m);
    HIPCHK(hipMemcpyAsync(dstart, seg_start, (g.m + 1ull) * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(drow, seg_row, g.m * 8ull, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dsrc, seg_src, g.m * 4ull, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dsrcs, srcs, g.m * sizeof(GatherSource), hipMemcpyHostToDevice, ctx->stream));
    uint64_t *keys = ctx->alloc<uint64_t>(std::max<uint64_t>(rows, 1));
    g.keep.push_back(keys);
    OutColSet cols = alloc_cols(ctx, v.cols, rows, g.keep);
    gather_segments(ctx->stream, dstart, n_seg, dsrc, drow, dsrcs, g.m, rows, keys, cols, ctx->tm());
    v.keys = keys;
    v.cols = as_in(cols);
    return v;
}

// Queues the group's partition passes and the readback of each join's
// largest leaf (pinned); returns without waiting.
void group_start(phj_ctx *ctx, BatchGroup &g, bool in_place, const uint32_t *leaf_rows) {
    const uint32_t m = g.m;
    HIPCHK(hipEventRecord(g.e2e.a, ctx->stream));
    uint32_t *r_off = ctx->pinned_io.get<uint32_t>(m + 1);
    uint32_t *s_off = ctx->pinned_io.get<uint32_t>(m + 1);
    r_off[0] = s_off[0] = 0;
    g.base.resize(m);
    g.end.resize(m);
    std::vector<uint32_t> jbits(m);
    uint32_t n_passes = 0;
    for (uint32_t j = 0; j < m; ++j) {
        r_off[j + 1] = r_off[j] + static_cast<uint32_t>(g.d[j].build.rows);
        s_off[j + 1] = s_off[j] + static_cast<uint32_t>(g.d[j].probe.rows);
        g.base[j] = s_off[j];
        g.end[j] = s_off[j + 1];
        jbits[j] = batch_bits(g.d[j].build.rows, leaf_rows);
        n_passes = std::max(n_passes, (jbits[j] + kDeviceMaxPassBits - 1) / kDeviceMaxPassBits);
    }
    std::vector<std::vector<uint32_t>> bits(n_passes, std::vector<uint32_t>(m, 0));
    for (uint32_t j = 0; j < m; ++j) {
        uint32_t left = jbits[j];
        for (uint32_t p = 0; p < n_passes; ++p) {
            bits[p][j] = std::min(left, kDeviceMaxPassBits);
            left -= bits[p][j];
        }
    }
    uint32_t *dr_off = ctx->alloc<uint32_t>(m + 1);
    g.keep.push_back(dr_off);
    uint32_t *ds_off = ctx->alloc<uint32_t>(m + 1);
    g.keep.push_back(ds_off);
    HIPCHK(hipMemcpyAsync(dr_off, r_off, (m + 1) * 4ull, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ds_off, s_off, (m + 1) * 4ull, hipMemcpyHostToDevice, ctx->stream));
    g.R = group_side(ctx, g, 0, in_place, r_off[m]);
    g.S = group_side(ctx, g, 1, in_place, s_off[m]);
    const phj_tuning tun = g.d[0].tuning;
    HIPCHK(hipEventRecord(g.ebs.a, ctx->stream));
    std::vector<uint32_t> leaf_join;
    LeafJoinArgs &a = g.a;
    a.r_cols = g.R.cols;
    a.s_cols = g.S.cols;
    if (n_passes > 0) {
        const Partitioned RP = radix_partition_joins(ctx, g.R, dr_off, m, bits, g.keep, radix_flags(tun), leaf_join);
        HIPCHK(hipEventRecord(g.ebs.b, ctx->stream));
        HIPCHK(hipEventRecord(g.eps.a, ctx->stream));
        std::vector<uint32_t> lj2;
        const Partitioned SP = radix_partition_joins(ctx, g.S, ds_off, m, bits, g.keep, radix_flags(tun), lj2);
        HIPCHK(hipEventRecord(g.eps.b, ctx->stream));
        a.r_keys = RP.keys;
        a.r_cols = as_in(RP.cols);
        a.r_off = RP.off;
        a.s_keys = SP.keys;
        a.s_cols = as_in(SP.cols);
        a.s_off = SP.off;
    } else {
        HIPCHK(hipEventRecord(g.ebs.b, ctx->stream));
        HIPCHK(hipEventRecord(g.eps.a, ctx->stream));
        HIPCHK(hipEventRecord(g.eps.b, ctx->stream));
        leaf_join.resize(m);
        for (uint32_t j = 0; j < m; ++j)
            leaf_join[j] = j;
        a.r_keys = g.R.keys;
        a.r_off = dr_off;
        a.s_keys = g.S.keys;
        a.s_off = ds_off;
    }
    g.n_leaves = static_cast<uint32_t>(leaf_join.size());
    uint32_t *lj_h = ctx->pinned_io.get<uint32_t>(g.n_leaves);
    std::copy(leaf_join.begin(), leaf_join.end(), lj_h);
    uint32_t *dleaf = ctx->alloc<uint32_t>(g.n_leaves);
    g.keep.push_back(dleaf);
    HIPCHK(hipMemcpyAsync(dleaf, lj_h, g.n_leaves * 4ull, hipMemcpyHostToDevice, ctx->stream));
    a.leaf_join = dleaf;
    uint32_t *dmax = ctx->alloc<uint32_t>(m);
    g.keep.push_back(dmax);
    launch_fill_u32(ctx->stream, dmax, m, 0);
    if (g.n_leaves)
        k_join_max_leaf<<<(g.n_leaves + 255) / 256, 256, 0, ctx->stream>>>(a.r_off, g.n_leaves, dleaf, dmax);
    g.join_max_h = ctx->pinned_io.get<uint32_t>(m);
    HIPCHK(hipMemcpyAsync(g.join_max_h, dmax, m * 4ull, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipEventRecord(g.ej.a, ctx->stream)); // the host waits for this point only
}

// Waits for the group's leaf sizes, then queues its fused leaf join and the
// readback of every join's output cursor and repeated-key flag.
void group_join(phj_ctx *ctx, BatchGroup &g) {
    HIPCHK(hipEventSynchronize(g.ej.a));
    const uint32_t m = g.m;
    uint32_t max_leaf = 0;
    for (uint32_t j = 0; j < m; ++j)
        max_leaf = std::max(max_leaf, g.join_max_h[j]);
    const LeafJoinPlan plan = plan_of(g.d[0].tuning);
    uint32_t lds_cap = leaf_join_lds_max_rows(78u * 1024u, plan.load_pct);
    if (max_leaf > lds_cap)
        lds_cap = leaf_join_lds_max_rows(150u * 1024u, plan.load_pct);
    g.lds_max = std::max<uint32_t>(std::min(max_leaf, lds_cap), 1);
    g.rerun.assign(m, 0);
    for (uint32_t j = 0; j < m; ++j)
        g.rerun[j] = g.join_max_h[j] > g.lds_max; // leaves beyond LDS: the plain path
    LeafJoinArgs &a = g.a;
    g.out = alloc_output(ctx, g.R.cols, g.S.cols, std::max<uint64_t>(g.S.rows, 1));
    Scratch scr(ctx);
    unsigned long long *cur = scr.get<unsigned long long>(m);
    uint64_t *dend = scr.get<uint64_t>(m);
    uint32_t *dflag = scr.get<uint32_t>(m);
    uint64_t *base_h = ctx->pinned_io.get<uint64_t>(m);
    uint64_t *end_h = ctx->pinned_io.get<uint64_t>(m);
    std::copy(g.base.begin(), g.base.end(), base_h);
    std::copy(g.end.begin(), g.end.end(), end_h);
    HIPCHK(hipMemcpyAsync(cur, base_h, m * 8ull, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dend, end_h, m * 8ull, hipMemcpyHostToDevice, ctx->stream));
    launch_fill_u32(ctx->stream, dflag, m, 0);
    a.n_leaves = g.n_leaves;
    a.lds_max_rows = g.lds_max;
    a.load_pct = plan.load_pct;
    a.chunk_rows = leaf_join_fused_chunk();
    a.join_cursor = cur;
    a.join_end = dend;
    a.join_flag = dflag;
    a.out = sink_of(g.out, nullptr);
    uint32_t *work_first = scr.get<uint32_t>(g.n_leaves + 1);
    uint32_t *scan_tmp = scr.get<uint32_t>(radix_scan_blocks(g.n_leaves + 1) + 1);
    const uint64_t max_work = leaf_max_work(g.S.rows, g.n_leaves, a.chunk_rows);
    uint32_t *item_leaf = scr.get<uint32_t>(max_work);
    a.work_first = work_first;
    a.item_leaf = item_leaf;
    leaf_work_plan(ctx->stream, a, work_first, scan_tmp, item_leaf, ctx->tm());
    leaf_join_fused(ctx->stream, a, max_work, ctx->tm());
    HIPCHK(hipEventRecord(g.ej.b, ctx->stream));
    HIPCHK(hipEventRecord(g.e2e.b, ctx->stream));
    g.cur_h = ctx->pinned_io.get<unsigned long long>(m);
    g.flag_h = ctx->pinned_io.get<uint32_t>(m);
    HIPCHK(hipMemcpyAsync(g.cur_h, cur, m * 8ull, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(g.flag_h, dflag, m * 4ull, hipMemcpyDeviceToHost, ctx->stream));
    // stream order keeps the partitions alive for the queued join
    g.keep.free_all();
}

// After the stream has finished the group: its results, one view of the
// shared output per join; joins to rerun are marked.
void group_collect(phj_ctx *ctx, BatchGroup &g, uint64_t peak) {
    const double t_bs = g.ebs.ms(), t_ps = g.eps.ms(), t_j = g.ej.ms(), t_e2e = g.e2e.ms();
    auto bufs = std::make_shared<std::vector<void *>>(g.out.bufs);
    g.out.bufs.clear();
    for (uint32_t j = 0; j < g.m; ++j) {
        phj_join_result *r = &g.res[j];
        std::memset(r, 0, sizeof(*r));
        if (g.flag_h[j] != 0)
            g.rerun[j] = 1;
        if (g.rerun[j])
            continue;
        OutBuffers view;
        view.keys = g.out.keys + g.base[j];
        view.left = g.out.left;
        view.right = g.out.right;
        for (uint32_t c = 0; c < g.out.left.n; ++c)
            view.left.ptr[c] = static_cast<uint8_t *>(g.out.left.ptr[c]) + g.base[j] * type_size(g.out.left.types[c]);
        for (uint32_t c = 0; c < g.out.right.n; ++c)
            view.right.ptr[c] = static_cast<uint8_t *>(g.out.right.ptr[c]) + g.base[j] * type_size(g.out.right.types[c]);
        fill_result_cols(r, view, g.cur_h[j] - g.base[j]);
        // Members of a group report the group's phase times.
        set_phase(r->build_shuffle, t_bs, g.R.rows);
        set_phase(r->probe_shuffle, t_ps, g.S.rows);
        set_phase(r->build, 0.0, g.R.rows);
        set_phase(r->probe, t_j, g.S.rows);
        r->e2e_wall_ms = t_e2e;
        r->peak_mem_bytes = peak;
        r->handle = ctx->next_handle++;
        ctx->shared[r->handle] = bufs;
    }
    if (bufs.use_count() == 1)
        for (void *b : *bufs)
            ctx->arena.free(b);
}

void run_phj_batch(phj_ctx *ctx, uint32_t n, const phj_join_desc *d, phj_join_result *res, const uint32_t *leaf_rows) {
    if (n > 0 && (d == nullptr || res == nullptr))
        invalid("null descriptors/results");
    for (uint32_t i = 0; i < n; ++i) {
        if (d[i].scheme != PHJ_SCHEME_PHJ_PURE)
            invalid("batched joins run PHJ-PURE");
        validate_rel(&d[i].build);
        validate_rel(&d[i].probe);
        validate_pass_bits(d[i].pass_bits, d[i].n_passes);
        std::memset(&res[i], 0, sizeof(res[i]));
    }
    // A join can be grouped when both sides are device relations (a group
    // reads them in place or gathers them; host relations take the plain
    // path, which stages them) and it asks for no other join path.
    auto groupable = [&](uint32_t i) {
        return d[i].build.on_device && d[i].probe.on_device && !(d[i].tuning.flags & PHJ_TUNE_TWO_PHASE_JOIN) &&
               !(d[i].tuning.flags & PHJ_TUNE_NARROW_PROBE) && d[i].build.rows <= 0xFFFFFFF0ull &&
               d[i].probe.rows <= 0xFFFFFFF0ull;
    };
    struct Span {
        uint32_t i, e;
        bool in_place;
    };
    std::vector<Span> spans;
    std::vector<uint32_t> plain;
    auto join_bytes = [&](uint32_t i) { return rel_bytes(d[i].build) + rel_bytes(d[i].probe); };
    for (uint32_t i = 0; i < n;) {
        if (!groupable(i)) {
            plain.push_back(i++);
            continue;
        }
        auto fits = [&](uint32_t e) {
            return e < n && groupable(e) && same_schema(d[i].build, d[e].build) && same_schema(d[i].probe, d[e].probe);
        };
        // joins lying back to back form a group as they are
        uint64_t bytes = join_bytes(i);
        uint32_t e = i + 1;
        while (fits(e) && follows(d[e - 1].build, d[e].build) && follows(d[e - 1].probe, d[e].probe) &&
               bytes + join_bytes(e) <= kBatchGroupBytes)
            bytes += join_bytes(e++);
        bool in_place = true;
        if (e == i + 1 && join_bytes(i) <= kBatchStageBytes) {
            // otherwise a run of small joins is gathered back to back
            uint32_t k = i + 1;
            bytes = join_bytes(i);
            while (fits(k) && join_bytes(k) <= kBatchStageBytes && bytes + join_bytes(k) <= kBatchGroupBytes)
                bytes += join_bytes(k++);
            if (k > i + 1) {
                e = k;
                in_place = false;
            }
        }
        spans.push_back(Span{i, e, in_place});
        i = e;
    }
    ctx->arena.reset_peak();
    std::vector<std::unique_ptr<BatchGroup>> groups;
    for (const Span &sp : spans) {
        auto g = std::make_unique<BatchGroup>(ctx);
        g->d = d + sp.i;
        g->res = res + sp.i;
        g->m = sp.e - sp.i;
        groups.push_back(std::move(g));
    }
    // pipeline: start(g + 1) is queued before the host waits inside join(g)
    if (!groups.empty())
        group_start(ctx, *groups[0], spans[0].in_place, leaf_rows);
    for (size_t g = 0; g < groups.size(); ++g) {
        if (g + 1 < groups.size())
            group_start(ctx, *groups[g + 1], spans[g + 1].in_place, leaf_rows);
        group_join(ctx, *groups[g]);
    }
    ctx->sync();
    const uint64_t peak = ctx->arena.peak();
    for (auto &g : groups)
        group_collect(ctx, *g, peak);
    ctx->pinned_io.reset();
    for (size_t g = 0; g < groups.size(); ++g)
        for (uint32_t j = 0; j < groups[g]->m; ++j)
            if (groups[g]->rerun[j]) {
                const phj_join_desc &x = groups[g]->d[j];
                run_phj_pure(ctx, x.build, x.probe, x.pass_bits, x.n_passes, x.tuning, &groups[g]->res[j]);
            }
    for (uint32_t i : plain)
        run_phj_pure(ctx, d[i].build, d[i].probe, d[i].pass_bits, d[i].n_passes, d[i].tuning, &res[i]);
}

__global__ void k_u64_to_u32(const uint64_t *in, uint32_t *out, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = static_cast<uint32_t>(in[i]);
}

__global__ void k_rows_to_refs(const uint32_t *rows, uint64_t n, uint32_t block_rows, phj_row_ref *out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t r = rows[i];
        phj_row_ref ref;
        ref.block_no = r == kInvalidRow ? PHJ_INVALID_BLOCK : r / block_rows;
        ref.row_no = r == kInvalidRow ? 0u : r % block_rows;
        out[i] = ref;
    }
}

void fill_partitioned(phj_ctx *ctx, const Partitioned &P, std::vector<void *> &keep, phj_partitioned *out) {
    std::memset(out, 0, sizeof(*out));
    uint64_t *off64 = ctx->alloc<uint64_t>(P.parts + 1);
    keep.push_back(off64);
    launch_u32_to_u64(ctx->stream, P.off, off64, P.parts + 1);
    out->rows = P.rows;
    out->parts = P.parts;
    out->keys = P.keys;
    out->offsets = off64;
    out->n_cols = P.cols.n;
    for (uint32_t c = 0; c < P.cols.n; ++c) {
        out->col_types[c] = P.cols.types[c];
        out->cols[c] = P.cols.ptr[c];
    }
    const uint64_t h = ctx->next_handle++;
    ctx->owned[h] = keep;
    keep.clear(); // owned by the handle now
    out->handle = h;
    ctx->sync();
}

GenParams make_gen_params(phj_ctx *ctx, const phj_gen_spec &s, Scratch &scr) {
    GenParams g{};
    g.seed = s.seed;
    g.build_rows = s.build_rows;
    g.build_total = s.build_rows;
    g.probe_rows = s.probe_rows;
    g.dist = s.distribution;
    g.match_ppm = s.match_ppm;
    if (s.build_n_cols > PHJ_MAX_COLS || s.probe_n_cols > PHJ_MAX_COLS)
        invalid("at most 8 payload columns");
    if (s.distribution > 2)
        invalid("unknown distribution");
    if (s.distribution == 1 && s.build_rows == 0 && s.probe_rows > 0)
        invalid("reference streams need build rows");
    g.n_build_cols = s.build_n_cols;
    g.n_probe_cols = s.probe_n_cols;
    for (uint32_t c = 0; c < PHJ_MAX_COLS; ++c) {
        g.build_types[c] = s.build_types[c];
        g.probe_types[c] = s.probe_types[c];
        if (c < s.build_n_cols)
            tsize(s.build_types[c]);
        if (c < s.probe_n_cols)
            tsize(s.probe_types[c]);
        g.b_bpay[c] = stream_base(s.seed, kStreamBuildPay + c);
        g.b_bpayhi[c] = stream_base(s.seed, kStreamBuildPayHi + c);
        g.b_ppay[c] = stream_base(s.seed, kStreamProbePay + c);
        g.b_ppayhi[c] = stream_base(s.seed, kStreamProbePayHi + c);
    }
    g.b_sel = stream_base(s.seed, kStreamProbeSelect);
    g.b_pick = stream_base(s.seed, kStreamProbePick);
    g.b_miss = stream_base(s.seed, kStreamProbeMiss);
    g.b_rank = stream_base(s.seed, kStreamZipfRank);
    const uint64_t kb = stream_base(s.seed, kStreamBuildKey);
    g.m1 = (rand64(kb, 0) | 1) & kMask63;
    g.m2 = (rand64(kb, 1) | 1) & kMask63;
    g.m3 = (rand64(kb, 2) | 1) & kMask63;
    if (s.distribution == 2) {
        // Zipf tables (53-bit CDF thresholds) and the duplicate prefix sums.
        auto thresholds = [](uint64_t n, double sx) {
            std::vector<double> cum(n);
            double acc = 0.0;
            for (uint64_t k = 1; k <= n; ++k) {
                acc += std::pow(static_cast<double>(k), -sx);
                cum[k - 1] = acc;
            }
            std::vector<uint64_t> t(n);
            const double scale = 9007199254740992.0;
            for (uint64_t k = 0; k < n; ++k)
                t[k] = static_cast<uint64_t>(cum[k] / acc * scale);
            if (n > 0)
                t[n - 1] = 1ull << 53;
            return t;
        };
        const uint64_t maxdup = s.zipf_build_max_dup == 0 ? 1 : s.zipf_build_max_dup;
        const std::vector<uint64_t> dup_t = thresholds(maxdup, s.zipf_build_s_milli / 1000.0);
        const std::vector<uint64_t> rank_t = thresholds(s.build_rows, s.zipf_probe_s_milli / 1000.0);
        std::vector<uint64_t> prefix(s.build_rows + 1, 0);
        const uint64_t bd = stream_base(s.seed, kStreamZipfDup);
        for (uint64_t d = 0; d < s.build_rows; ++d) {
            const uint64_t u = rand64(bd, d) >> 11;
            const uint64_t k = static_cast<uint64_t>(std::upper_bound(dup_t.begin(), dup_t.end(), u) - dup_t.begin());
            prefix[d + 1] = prefix[d] + 1 + k;
        }
        g.build_total = prefix[s.build_rows];
        uint64_t *dp = scr.get<uint64_t>(prefix.size());
        uint64_t *dr = scr.get<uint64_t>(rank_t.size());
        HIPCHK(hipMemcpyAsync(dp, prefix.data(), prefix.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        if (!rank_t.empty())
            HIPCHK(hipMemcpyAsync(dr, rank_t.data(), rank_t.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        g.prefix = dp;
        g.rank_t = dr;
        g.n_rank = static_cast<uint32_t>(rank_t.size());
        const uint64_t sb = stream_base(s.seed, kStreamShuffle);
        unsigned m = 0;
        while ((1ull << m) < g.build_total && m < 63)
            ++m;
        g.sh_mask = m == 0 ? 0 : ((1ull << m) - 1);
        g.sh_r = (m + 1) / 2 == 0 ? 1 : (m + 1) / 2;
        g.sh_o1 = rand64(sb, 10) | 1;
        g.sh_o2 = rand64(sb, 11) | 1;
        g.sh_c1 = rand64(sb, 12);
        g.sh_i1 = inverse_odd(g.sh_o1);
        g.sh_i2 = inverse_odd(g.sh_o2);
    }
    return g;
}

void gen_side(phj_ctx *ctx, const GenParams &g, int side, uint64_t rows, uint32_t ncols, const uint32_t *types,
              phj_relation *rel) {
    std::memset(rel, 0, sizeof(*rel));
    rel->rows = rows;
    rel->on_device = 1;
    rel->n_cols = ncols;
    uint64_t *keys = ctx->alloc<uint64_t>(std::max<uint64_t>(rows, 1));
    rel->keys = keys;
    OutColSet oc{};
    oc.n = ncols;
    for (uint32_t c = 0; c < ncols; ++c) {
        oc.types[c] = types[c];
        oc.ptr[c] = ctx->arena.alloc(std::max<uint64_t>(rows * type_size(types[c]), 1));
        rel->cols[c].type = types[c];
        rel->cols[c].data = oc.ptr[c];
    }
    gen_fill(ctx->stream, g, side, rows, keys, oc);
}

ColSet colset_of_result(const phj_join_result *r, bool left) {
    ColSet c{};
    c.n = left ? r->n_left : r->n_right;
    for (uint32_t i = 0; i < c.n; ++i) {
        c.types[i] = left ? r->left_types[i] : r->right_types[i];
        c.ptr[i] = left ? r->left[i] : r->right[i];
    }
    return c;
}

// Returns a result's device memory: its own blocks, or its share of a batch
// group's output (freed with the group's last result).
void release_result(phj_ctx *ctx, phj_join_result &r) {
    if (r.handle == 0)
        return;
    auto it = ctx->owned.find(r.handle);
    if (it != ctx->owned.end()) {
        for (void *b : it->second)
            ctx->arena.free(b);
        ctx->owned.erase(it);
    }
    auto sh = ctx->shared.find(r.handle);
    if (sh != ctx->shared.end()) {
        const std::shared_ptr<std::vector<void *>> bufs = sh->second;
        ctx->shared.erase(sh);
        if (bufs.use_count() == 1)
            for (void *b : *bufs)
                ctx->arena.free(b);
    }
    r.handle = 0;
}

template <class F> phj_status guard(phj_ctx *ctx, F &&f) {
    if (ctx == nullptr)
        return PHJ_ERR_INVALID_ARGUMENT;
    try {
        HIPCHK(hipSetDevice(ctx->device));
        f();
        ctx->err.clear();
        return PHJ_OK;
    } catch (const PhjError &e) {
        ctx->err = e.what();
        return e.code;
    } catch (const ArenaError &e) {
        ctx->err = e.what();
        return e.kind == ArenaError::kOutOfMemory ? PHJ_ERR_OUT_OF_MEMORY : PHJ_ERR_INVALID_ARGUMENT;
    } catch (const LaunchTooLarge &e) {
        ctx->err = e.what();
        return PHJ_ERR_UNSUPPORTED;
    } catch (const std::bad_alloc &) {
        ctx->err = "host allocation failed";
        return PHJ_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return PHJ_ERR_HIP;
    }
}

} // namespace

extern "C" {

uint32_t phj_abi_version(void) { return PHJ_ABI_VERSION; }

phj_status phj_ctx_create(int device, phj_ctx **out) {
    if (out == nullptr)
        return PHJ_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return PHJ_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n)
        return PHJ_ERR_INVALID_ARGUMENT;
    auto ctx = std::make_unique<phj_ctx>();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
        return PHJ_ERR_HIP;
    if (scratch_address(device, &ctx->scratch) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&ctx->pinned), 64, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->note_event, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        if (ctx->pinned)
            (void)hipHostFree(ctx->pinned);
        (void)hipStreamDestroy(ctx->stream);
        return PHJ_ERR_HIP;
    }
    ctx->arena.bind(ctx->stream);
    *out = ctx.release();
    return PHJ_OK;
}

phj_status phj_ctx_destroy(phj_ctx *ctx) {
    if (ctx == nullptr)
        return PHJ_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->owned)
        for (void *p : kv.second)
            ctx->arena.free(p);
    ctx->owned.clear();
    ctx->watch_release();
    ctx->arena.release_cached();
    (void)hipEventDestroy(ctx->note_event);
    (void)hipHostFree(ctx->pinned);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PHJ_OK;
}

const char *phj_last_error(const phj_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

phj_status phj_synchronize(phj_ctx *ctx) {
    return guard(ctx, [&] { ctx->sync(); });
}

phj_status phj_dev_alloc(phj_ctx *ctx, uint64_t bytes, void **out) {
    return guard(ctx, [&] {
        if (out == nullptr)
            invalid("null output pointer");
        *out = ctx->arena.alloc(bytes);
    });
}

phj_status phj_placement_stats(phj_ctx *ctx, uint64_t *marked_blocks, uint64_t *marked_bytes) {
    return guard(ctx, [&] {
        if (marked_blocks)
            *marked_blocks = ctx->watch_marked;
        if (marked_bytes)
            *marked_bytes = ctx->arena.slow_bytes();
    });
}

phj_status phj_dev_free(phj_ctx *ctx, void *ptr) {
    return guard(ctx, [&] { ctx->arena.free(ptr); });
}

phj_status phj_memcpy_h2d(phj_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    return guard(ctx, [&] {
        if (bytes)
            HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        ctx->sync();
    });
}

phj_status phj_memcpy_d2h(phj_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    return guard(ctx, [&] {
        if (bytes)
            HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
    });
}

phj_status phj_mem_stats(phj_ctx *ctx, uint64_t *current_bytes, uint64_t *peak_bytes) {
    return guard(ctx, [&] {
        if (current_bytes)
            *current_bytes = ctx->arena.current();
        if (peak_bytes)
            *peak_bytes = ctx->arena.peak();
    });
}

phj_status phj_reset_peak(phj_ctx *ctx) {
    return guard(ctx, [&] { ctx->arena.reset_peak(); });
}

phj_status phj_set_mem_limit(phj_ctx *ctx, uint64_t bytes) {
    return guard(ctx, [&] { ctx->arena.set_limit(bytes); });
}

phj_status phj_int_hash64_batch(phj_ctx *ctx, const uint64_t *in, uint64_t n, uint64_t *out, int on_device) {
    return guard(ctx, [&] {
        if (n && (in == nullptr || out == nullptr))
            invalid("null buffer");
        if (on_device) {
            launch_hash_batch(ctx->stream, in, n, out, ctx->tm());
            ctx->sync();
            return;
        }
        Scratch scr(ctx);
        uint64_t *d = scr.get<uint64_t>(n);
        if (n)
            HIPCHK(hipMemcpyAsync(d, in, n * 8, hipMemcpyHostToDevice, ctx->stream));
        launch_hash_batch(ctx->stream, d, n, d, ctx->tm());
        if (n)
            HIPCHK(hipMemcpyAsync(out, d, n * 8, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
    });
}

phj_status phj_radix_shuffle(phj_ctx *ctx, const phj_relation *input, const uint8_t *pass_bits, uint32_t n_passes,
                             phj_partitioned *out) {
    return guard(ctx, [&] {
        validate_rel(input);
        if (out == nullptr)
            invalid("null output");
        const uint32_t total = validate_pass_bits(pass_bits, n_passes);
        Scratch scr(ctx);
        const DevRel R = stage_relation(ctx, scr, *input);
        Held keep(ctx);
        const Partitioned P = radix_partition(ctx, R, nullptr, 1, 0, device_passes(total), keep, "shuffle");
        fill_partitioned(ctx, P, keep, out);
    });
}

// radixShuffle(input, out_schema, in_col_index, cfg, threads)
// (src/RadixPartition.h:187-193, RadixPartition.cpp:546-566): the scatter
// moves input column in_col_index[c] into output column c. The column map is
// applied to the relation's column pointers, so the device passes are those
// of phj_radix_shuffle; a column named twice is moved twice.
phj_status phj_radix_shuffle_project(phj_ctx *ctx, const phj_relation *input, const uint32_t *out_types,
                                     const uint32_t *in_col_index, uint32_t n_out_cols, const uint8_t *pass_bits,
                                     uint32_t n_passes, phj_partitioned *out) {
    return guard(ctx, [&] {
        validate_rel(input);
        if (out == nullptr)
            invalid("null output");
        const uint32_t total = validate_pass_bits(pass_bits, n_passes);
        if (n_out_cols > PHJ_MAX_COLS)
            invalid("at most 8 payload columns");
        if (n_out_cols > 0 && (out_types == nullptr || in_col_index == nullptr))
            invalid("radixShuffle: null out_schema or in_col_index");
        phj_relation proj = *input;
        proj.n_cols = n_out_cols;
        for (uint32_t c = 0; c < n_out_cols; ++c) {
            const uint32_t i = in_col_index[c];
            if (i >= input->n_cols)
                invalid("radixShuffle: in_col_index[" + std::to_string(c) + "] = " + std::to_string(i) +
                        " names no input column");
            // The reference moves the input column's bytes at the output
            // type's width; widths that differ would read past or split its
            // values, so the types must agree.
            if (out_types[c] != input->cols[i].type)
                invalid("radixShuffle: out_schema type of column " + std::to_string(c) +
                        " differs from the type of input column " + std::to_string(i));
            proj.cols[c] = input->cols[i];
        }
        Scratch scr(ctx);
        const DevRel R = stage_relation(ctx, scr, proj);
        Held keep(ctx);
        const Partitioned P = radix_partition(ctx, R, nullptr, 1, 0, device_passes(total), keep, "shuffle");
        fill_partitioned(ctx, P, keep, out);
    });
}

phj_status phj_refine_to_leaves(phj_ctx *ctx, const phj_partitioned *pass1, const uint8_t *pass_bits, uint32_t n_passes,
                                phj_partitioned *out) {
    return guard(ctx, [&] {
        if (pass1 == nullptr || out == nullptr)
            invalid("null partitioned input/output");
        if (pass_bits == nullptr || n_passes == 0)
            invalid("refineToLeaves: empty pass_bits"); // src/RadixPartition.cpp:425-426
        const uint32_t total = validate_pass_bits(pass_bits, n_passes);
        if (pass1->parts != (1ull << pass_bits[0]))
            invalid("input partition count must equal 2^pass_bits[0]");
        Scratch scr(ctx);
        DevRel in;
        in.rows = pass1->rows;
        in.keys = pass1->keys;
        in.cols.n = pass1->n_cols;
        for (uint32_t c = 0; c < pass1->n_cols; ++c) {
            in.cols.types[c] = pass1->col_types[c];
            in.cols.ptr[c] = pass1->cols[c];
        }
        uint32_t *off32 = scr.get<uint32_t>(pass1->parts + 1);
        k_u64_to_u32<<<static_cast<uint32_t>((pass1->parts + 1 + 255) / 256), 256, 0, ctx->stream>>>(pass1->offsets, off32,
                                                                                                     pass1->parts + 1);
        Held keep(ctx);
        Partitioned P;
        if (n_passes == 1) {
            // One pass: the pass-1 buckets are the leaves (src/RadixPartition.cpp:437-447); copy them.
            P.rows = in.rows;
            P.parts = pass1->parts;
            P.keys = ctx->alloc<uint64_t>(std::max<uint64_t>(in.rows, 1));
            keep.push_back(P.keys);
            launch_copy(ctx->stream, P.keys, in.keys, in.rows * 8, nullptr);
            P.cols = alloc_cols(ctx, in.cols, in.rows, keep);
            for (uint32_t c = 0; c < in.cols.n; ++c)
                launch_copy(ctx->stream, P.cols.ptr[c], in.cols.ptr[c], in.rows * type_size(in.cols.types[c]), nullptr);
            P.off = ctx->alloc<uint32_t>(P.parts + 1);
            keep.push_back(P.off);
            launch_copy(ctx->stream, P.off, off32, (P.parts + 1) * 4, nullptr);
        } else {
            const std::vector<uint32_t> passes = device_passes(total - pass_bits[0]);
            P = radix_partition(ctx, in, off32, static_cast<uint32_t>(pass1->parts), pass_bits[0], passes, keep, "refine");
        }
        fill_partitioned(ctx, P, keep, out);
    });
}

phj_status phj_partitioned_release(phj_ctx *ctx, phj_partitioned *p) {
    return guard(ctx, [&] {
        if (p == nullptr || p->handle == 0)
            return;
        auto it = ctx->owned.find(p->handle);
        if (it != ctx->owned.end()) {
            for (void *b : it->second)
                ctx->arena.free(b);
            ctx->owned.erase(it);
        }
        std::memset(p, 0, sizeof(*p));
    });
}

phj_status phj_hashtable_build_find(phj_ctx *ctx, const uint64_t *build_keys, uint64_t n_build, const uint64_t *probe_keys,
                                    uint64_t n_probe, uint32_t block_rows, int on_device, phj_row_ref *heads,
                                    phj_row_ref *next) {
    return guard(ctx, [&] {
        if (block_rows == 0)
            invalid("block_rows must be positive");
        if (n_build > 0xFFFFFFFEull)
            invalid("at most 2^32-2 build rows");
        Scratch scr(ctx);
        const uint64_t *bk = build_keys, *pk = probe_keys;
        if (!on_device) {
            uint64_t *b = scr.get<uint64_t>(n_build), *p = scr.get<uint64_t>(n_probe);
            if (n_build)
                HIPCHK(hipMemcpyAsync(b, build_keys, n_build * 8, hipMemcpyHostToDevice, ctx->stream));
            if (n_probe)
                HIPCHK(hipMemcpyAsync(p, probe_keys, n_probe * 8, hipMemcpyHostToDevice, ctx->stream));
            bk = b;
            pk = p;
        }
        uint64_t cap = 2; // src/HashTable.h:72-77 reserve: bit_ceil(2 * rows)
        while (cap < 2 * n_build)
            cap <<= 1;
        std::vector<uint64_t> off = {0, cap};
        uint64_t *d_off = scr.get<uint64_t>(2);
        HIPCHK(hipMemcpyAsync(d_off, off.data(), 16, hipMemcpyHostToDevice, ctx->stream));
        GlobalHT ht{};
        ht.cells = scr.get<Cell16>(cap);
        ht.cell_off = d_off;
        ht.next = scr.get<uint32_t>(n_build);
        ht.side_head = scr.get<uint32_t>(1);
        ht.n_tables = 1;
        ht.block_rows = block_rows;
        ht_init(ctx->stream, ht, cap, n_build, ctx->tm());
        HtRows hr{};
        hr.keys = bk;
        hr.rows = n_build;
        ht_build(ctx->stream, ht, kHtSingle, hr, n_build, ctx->tm(), "ht_build");
        uint32_t *hrows = scr.get<uint32_t>(n_probe);
        ht_find(ctx->stream, ht, pk, n_probe, nullptr, hrows, ctx->tm());
        phj_row_ref *dh = on_device ? heads : scr.get<phj_row_ref>(n_probe);
        phj_row_ref *dn = on_device ? next : scr.get<phj_row_ref>(n_build);
        if (n_probe)
            k_rows_to_refs<<<static_cast<uint32_t>((n_probe + 255) / 256), 256, 0, ctx->stream>>>(hrows, n_probe, block_rows, dh);
        if (n_build)
            k_rows_to_refs<<<static_cast<uint32_t>((n_build + 255) / 256), 256, 0, ctx->stream>>>(ht.next, n_build, block_rows,
                                                                                                  dn);
        if (!on_device) {
            if (n_probe && heads)
                HIPCHK(hipMemcpyAsync(heads, dh, n_probe * 8, hipMemcpyDeviceToHost, ctx->stream));
            if (n_build && next)
                HIPCHK(hipMemcpyAsync(next, dn, n_build * 8, hipMemcpyDeviceToHost, ctx->stream));
        }
        ctx->sync();
    });
}

phj_status phj_hashtable_build_find_noalloc(phj_ctx *ctx, const uint64_t *build_keys, uint64_t n_build,
                                            const uint64_t *probe_keys, uint64_t n_probe, uint32_t block_rows,
                                            phj_row_ref *heads, phj_row_ref *next) {
    return guard(ctx, [&] {
        if (block_rows == 0)
            invalid("block_rows must be positive");
        if ((n_build && (build_keys == nullptr || next == nullptr)) || (n_probe && (probe_keys == nullptr || heads == nullptr)))
            invalid("null device buffer");
        if (n_build > kNoAllocMaxBuild)
            throw PhjError(PHJ_ERR_CAPACITY, "phj_hashtable_build_find_noalloc: " + std::to_string(n_build) +
                                                 " build rows exceed its scratch (at most " +
                                                 std::to_string(kNoAllocMaxBuild) + ")");
        unsigned char *base = static_cast<unsigned char *>(ctx->scratch);
        uint64_t cap = 2; // src/HashTable.h:72-77 reserve: bit_ceil(2 * rows)
        while (cap < 2 * n_build)
            cap <<= 1;
        uint64_t *words = reinterpret_cast<uint64_t *>(base + kNoAllocWordsOff);
        // cell offsets {0, cap} and the table's side head, filled by kernels
        launch_fill_u64(ctx->stream, words, 1, 0);
        launch_fill_u64(ctx->stream, words + 1, 1, cap);
        GlobalHT ht{};
        ht.cells = reinterpret_cast<Cell16 *>(base + kNoAllocCellsOff);
        ht.cell_off = words;
        ht.next = reinterpret_cast<uint32_t *>(base + kNoAllocNextOff);
        ht.side_head = reinterpret_cast<uint32_t *>(words + 2);
        ht.n_tables = 1;
        ht.block_rows = block_rows;
        ht_init(ctx->stream, ht, cap, n_build, ctx->tm());
        HtRows hr{};
        hr.keys = build_keys;
        hr.rows = n_build;
        ht_build(ctx->stream, ht, kHtSingle, hr, n_build, ctx->tm(), "ht_build");
        uint32_t *rows = reinterpret_cast<uint32_t *>(base + kNoAllocRowsOff);
        for (uint64_t a = 0; a < n_probe; a += kNoAllocProbeChunk) {
            const uint64_t m = std::min(kNoAllocProbeChunk, n_probe - a);
            ht_find(ctx->stream, ht, probe_keys + a, m, nullptr, rows, ctx->tm());
            k_rows_to_refs<<<static_cast<uint32_t>((m + 255) / 256), 256, 0, ctx->stream>>>(rows, m, block_rows, heads + a);
        }
        if (n_build)
            k_rows_to_refs<<<static_cast<uint32_t>((n_build + 255) / 256), 256, 0, ctx->stream>>>(ht.next, n_build,
                                                                                                  block_rows, next);
        ctx->sync();
    });
}

phj_status phj_join(phj_ctx *ctx, const phj_join_desc *d, phj_join_result *result) {
    return guard(ctx, [&] {
        if (d == nullptr || result == nullptr)
            invalid("null descriptor/result");
        switch (d->scheme) {
        case PHJ_SCHEME_CHJ: run_chj(ctx, d->build, d->probe, result); break;
        case PHJ_SCHEME_PHJ_PURE:
            run_phj_pure(ctx, d->build, d->probe, d->pass_bits, d->n_passes, d->tuning, result);
            break;
        case PHJ_SCHEME_PHJ:
            run_phj_bep(ctx, d->build, d->probe, d->pass_bits, d->n_passes, d->bep_budget_bytes, d->tuning, result);
            break;
        default: invalid("unknown scheme");
        }
    });
}

phj_status phj_build_register(phj_ctx *ctx, const phj_relation *build, const uint8_t *pass_bits, uint32_t n_passes,
                              const phj_tuning *tuning, phj_resident **out) {
    return guard(ctx, [&] {
        if (build == nullptr || out == nullptr)
            invalid("null argument");
        *out = nullptr;
        const phj_tuning t = tuning ? *tuning : phj_tuning{};
        auto r = std::make_unique<phj_resident>();
        try {
            register_build(ctx, *build, pass_bits, n_passes, t, r.get());
        } catch (...) {
            for (void *b : r->bufs)
                ctx->arena.free(b);
            throw;
        }
        *out = r.release();
    });
}

phj_status phj_probe_registered(phj_ctx *ctx, const phj_resident *build, const phj_relation *probe,
                                const phj_tuning *tuning, phj_join_result *result) {
    return guard(ctx, [&] {
        if (build == nullptr || probe == nullptr || result == nullptr)
            invalid("null argument");
        if (build->owner->device != ctx->device)
            invalid("resident build side lives on another device");
        const phj_tuning t = tuning ? *tuning : phj_tuning{};
        auto *r = const_cast<phj_resident *>(build);
        struct InFlight {
            phj_resident *r;
            explicit InFlight(phj_resident *x) : r(x) { r->inflight.fetch_add(1, std::memory_order_acq_rel); }
            ~InFlight() { r->inflight.fetch_sub(1, std::memory_order_acq_rel); }
        } counted(r);
        probe_resident(ctx, *build, *probe, t, result);
    });
}

phj_status phj_build_release(phj_ctx *ctx, phj_resident *build) {
    return guard(ctx, [&] {
        if (build == nullptr)
            return;
        if (build->owner != ctx)
            invalid("release a resident build side through the context that registered it");
        if (build->inflight.load(std::memory_order_acquire) != 0)
            invalid("resident build side has probes in flight; release it after they return");
        for (void *b : build->bufs)
            ctx->arena.free(b);
        delete build;
    });
}

phj_status phj_build_bytes(const phj_resident *build, uint64_t *bytes) {
    if (build == nullptr || bytes == nullptr)
        return PHJ_ERR_INVALID_ARGUMENT;
    *bytes = build->bytes;
    return PHJ_OK;
}

phj_status phj_join_batch(phj_ctx *ctx, uint32_t n, const phj_join_desc *descs, phj_join_result *results) {
    return guard(ctx, [&] { run_phj_batch(ctx, n, descs, results, ctx->batch_leaf_rows); });
}

phj_status phj_batch_autotune(phj_ctx *ctx, uint32_t n, const phj_join_desc *descs, uint32_t *leaf_rows_out,
                              double *class_ms_out) {
    return guard(ctx, [&] {
        if (n > 0 && descs == nullptr)
            invalid("null descriptors");
        // joins by size class
        std::vector<std::vector<phj_join_desc>> by_class(kBatchClasses);
        for (uint32_t i = 0; i < n; ++i)
            by_class[size_class(descs[i].build.rows)].push_back(descs[i]);
        const uint32_t candidates[] = {512, 1024, 2048, 4096};
        std::vector<double> best_ms(kBatchClasses, 0.0);
        for (uint32_t c = 0; c < kBatchClasses; ++c) {
            auto &v = by_class[c];
            if (v.empty())
                continue;
            std::vector<phj_join_result> res(v.size());
            uint32_t table[kBatchClasses];
            std::copy(std::begin(ctx->batch_leaf_rows), std::end(ctx->batch_leaf_rows), table);
            double best = 1e300;
            uint32_t pick = table[c];
            for (uint32_t cand : candidates) {
                if ((1ull << c) <= cand && cand != candidates[0])
                    continue; // joins of this class fit one leaf at every candidate: nothing to tell apart
                table[c] = cand;
                double t_min = 1e300;
                for (int rep = 0; rep < 3; ++rep) { // the first run warms the pool
                    ctx->sync();
                    const auto t0 = std::chrono::steady_clock::now();
                    run_phj_batch(ctx, static_cast<uint32_t>(v.size()), v.data(), res.data(), table);
                    ctx->sync();
                    const double ms =
                        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                    for (auto &r : res)
                        release_result(ctx, r);
                    if (rep > 0)
                        t_min = std::min(t_min, ms);
                }
                if (t_min < best) {
                    best = t_min;
                    pick = cand;
                }
            }
            ctx->batch_leaf_rows[c] = pick;
            best_ms[c] = best;
        }
        if (leaf_rows_out)
            std::copy(std::begin(ctx->batch_leaf_rows), std::end(ctx->batch_leaf_rows), leaf_rows_out);
        if (class_ms_out)
            std::copy(best_ms.begin(), best_ms.end(), class_ms_out);
    });
}

phj_status phj_batch_set_leaf_rows(phj_ctx *ctx, const uint32_t *leaf_rows) {
    return guard(ctx, [&] {
        if (leaf_rows == nullptr)
            invalid("null table");
        for (uint32_t c = 0; c < kBatchClasses; ++c)
            if (leaf_rows[c] == 0)
                invalid("leaf rows must be positive");
        std::copy(leaf_rows, leaf_rows + kBatchClasses, ctx->batch_leaf_rows);
    });
}

phj_status phj_run_chj(phj_ctx *ctx, const phj_relation *build, const phj_relation *probe, uint32_t,
                       phj_join_result *result) {
    return guard(ctx, [&] {
        if (!build || !probe || !result)
            invalid("null argument");
        run_chj(ctx, *build, *probe, result);
    });
}

phj_status phj_run_phj_pure(phj_ctx *ctx, const phj_relation *build, const phj_relation *probe, const uint8_t *pass_bits,
                            uint32_t n_passes, uint32_t, phj_join_result *result) {
    return guard(ctx, [&] {
        if (!build || !probe || !result)
            invalid("null argument");
        phj_tuning t{};
        run_phj_pure(ctx, *build, *probe, pass_bits, n_passes, t, result);
    });
}

phj_status phj_run_phj(phj_ctx *ctx, const phj_relation *build, const phj_relation *probe, const uint8_t *pass_bits,
                       uint32_t n_passes, uint32_t, uint64_t bep_budget_mib, phj_join_result *result) {
    return guard(ctx, [&] {
        if (!build || !probe || !result)
            invalid("null argument");
        phj_tuning t{};
        run_phj_bep(ctx, *build, *probe, pass_bits, n_passes, bep_budget_mib << 20, t, result);
    });
}

phj_status phj_result_release(phj_ctx *ctx, phj_join_result *result) {
    return guard(ctx, [&] {
        if (result == nullptr || result->handle == 0)
            return;
        release_result(ctx, *result);
        std::memset(result, 0, sizeof(*result));
    });
}

phj_status phj_result_copy_to_host(phj_ctx *ctx, const phj_join_result *r, uint64_t *keys, void *const *left,
                                   void *const *right) {
    return guard(ctx, [&] {
        if (r == nullptr)
            invalid("null result");
        const uint64_t n = r->out_rows;
        if (n == 0)
            return;
        if (keys)
            HIPCHK(hipMemcpyAsync(keys, r->keys, n * 8, hipMemcpyDeviceToHost, ctx->stream));
        for (uint32_t c = 0; c < r->n_left; ++c)
            if (left && left[c])
                HIPCHK(hipMemcpyAsync(left[c], r->left[c], n * type_size(r->left_types[c]), hipMemcpyDeviceToHost,
                                      ctx->stream));
        for (uint32_t c = 0; c < r->n_right; ++c)
            if (right && right[c])
                HIPCHK(hipMemcpyAsync(right[c], r->right[c], n * type_size(r->right_types[c]), hipMemcpyDeviceToHost,
                                      ctx->stream));
        ctx->sync();
    });
}

phj_status phj_result_checksum(phj_ctx *ctx, const phj_join_result *r, phj_checksum *out) {
    return guard(ctx, [&] {
        if (r == nullptr || out == nullptr)
            invalid("null argument");
        Scratch scr(ctx);
        unsigned long long *acc = scr.get<unsigned long long>(3);
        checksum_columns(ctx->stream, r->out_rows, r->keys, colset_of_result(r, true), colset_of_result(r, false), acc,
                         ctx->tm());
        unsigned long long h[3];
        HIPCHK(hipMemcpyAsync(h, acc, 24, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        out->rows = h[0];
        out->lane0 = h[1];
        out->lane1 = h[2];
    });
}

phj_status phj_gen_relations(phj_ctx *ctx, const phj_gen_spec *spec, phj_relation *build, phj_relation *probe) {
    return guard(ctx, [&] {
        if (spec == nullptr || build == nullptr || probe == nullptr)
            invalid("null argument");
        if (spec->probe_rows > 0xFFFFFFFEull || spec->build_rows > 0xFFFFFFFEull)
            invalid("relations are limited to 2^32-2 rows");
        Scratch scr(ctx);
        const GenParams g = make_gen_params(ctx, *spec, scr);
        uint32_t bt[PHJ_MAX_COLS], pt[PHJ_MAX_COLS];
        for (uint32_t c = 0; c < PHJ_MAX_COLS; ++c) {
            bt[c] = spec->build_types[c];
            pt[c] = spec->probe_types[c];
        }
        gen_side(ctx, g, 0, g.build_total, spec->build_n_cols, bt, build);
        gen_side(ctx, g, 1, spec->probe_rows, spec->probe_n_cols, pt, probe);
        ctx->sync();
    });
}

phj_status phj_gen_batch(phj_ctx *ctx, uint32_t n, const phj_gen_spec *specs, phj_relation *build, phj_relation *probe,
                         uint64_t *build_off, uint64_t *probe_off) {
    return guard(ctx, [&] {
        if (n == 0 || specs == nullptr || build == nullptr || probe == nullptr || build_off == nullptr ||
            probe_off == nullptr)
            invalid("null argument");
        Scratch scr(ctx);
        std::vector<GenParams> g;
        g.reserve(n);
        build_off[0] = probe_off[0] = 0;
        for (uint32_t i = 0; i < n; ++i) {
            g.push_back(make_gen_params(ctx, specs[i], scr));
            if (specs[i].build_n_cols != specs[0].build_n_cols || specs[i].probe_n_cols != specs[0].probe_n_cols ||
                std::memcmp(specs[i].build_types, specs[0].build_types, PHJ_MAX_COLS) != 0 ||
                std::memcmp(specs[i].probe_types, specs[0].probe_types, PHJ_MAX_COLS) != 0)
                invalid("the joins of a generated batch share one schema");
            build_off[i + 1] = build_off[i] + g[i].build_total;
            probe_off[i + 1] = probe_off[i] + specs[i].probe_rows;
        }
        if (build_off[n] > 0xFFFFFFFEull * 64 || probe_off[n] > 0xFFFFFFFEull * 64)
            invalid("batch too large");
        uint32_t bt[PHJ_MAX_COLS], pt[PHJ_MAX_COLS];
        for (uint32_t c = 0; c < PHJ_MAX_COLS; ++c) {
            bt[c] = specs[0].build_types[c];
            pt[c] = specs[0].probe_types[c];
        }
        auto fill = [&](int side, const uint64_t *off, uint32_t ncols, const uint32_t *types, phj_relation *rel) {
            std::memset(rel, 0, sizeof(*rel));
            const uint64_t rows = off[n];
            rel->rows = rows;
            rel->on_device = 1;
            rel->n_cols = ncols;
            uint64_t *keys = ctx->alloc<uint64_t>(std::max<uint64_t>(rows, 1));
            rel->keys = keys;
            OutColSet oc{};
            oc.n = ncols;
            for (uint32_t c = 0; c < ncols; ++c) {
                oc.types[c] = types[c];
                oc.ptr[c] = ctx->arena.alloc(std::max<uint64_t>(rows * type_size(types[c]), 1));
                rel->cols[c].type = types[c];
                rel->cols[c].data = oc.ptr[c];
            }
            for (uint32_t i = 0; i < n; ++i) {
                OutColSet part = oc;
                for (uint32_t c = 0; c < ncols; ++c)
                    part.ptr[c] = static_cast<uint8_t *>(oc.ptr[c]) + off[i] * type_size(types[c]);
                gen_fill(ctx->stream, g[i], side, off[i + 1] - off[i], keys + off[i], part);
            }
        };
        fill(0, build_off, specs[0].build_n_cols, bt, build);
        fill(1, probe_off, specs[0].probe_n_cols, pt, probe);
        ctx->sync();
    });
}

phj_status phj_relation_release(phj_ctx *ctx, phj_relation *rel) {
    return guard(ctx, [&] {
        if (rel == nullptr)
            return;
        ctx->arena.free(const_cast<uint64_t *>(rel->keys));
        for (uint32_t c = 0; c < rel->n_cols; ++c)
            ctx->arena.free(const_cast<void *>(rel->cols[c].data));
        std::memset(rel, 0, sizeof(*rel));
    });
}

phj_status phj_gen_expected_checksum(phj_ctx *ctx, const phj_gen_spec *spec, phj_checksum *out) {
    return guard(ctx, [&] {
        if (spec == nullptr || out == nullptr)
            invalid("null argument");
        Scratch scr(ctx);
        const GenParams g = make_gen_params(ctx, *spec, scr);
        unsigned long long *acc = scr.get<unsigned long long>(3);
        gen_expected_checksum(ctx->stream, g, acc);
        unsigned long long h[3];
        HIPCHK(hipMemcpyAsync(h, acc, 24, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        out->rows = h[0];
        out->lane0 = h[1];
        out->lane1 = h[2];
    });
}

phj_status phj_set_kernel_timing(phj_ctx *ctx, int enabled) {
    return guard(ctx, [&] {
        ctx->sync();
        ctx->timing = enabled != 0;
    });
}

phj_status phj_kernel_times_reset(phj_ctx *ctx) {
    return guard(ctx, [&] {
        ctx->sync();
        ctx->timer.reset();
    });
}

phj_status phj_kernel_times(phj_ctx *ctx, const char *kernel, double *total_ms, uint64_t *launches) {
    return guard(ctx, [&] {
        if (kernel == nullptr)
            invalid("null kernel name");
        ctx->sync();
        double ms = 0;
        uint64_t n = 0;
        ctx->timer.get_total(kernel, ms, n);
        if (total_ms)
            *total_ms = ms;
        if (launches)
            *launches = n;
    });
}

phj_status phj_kernel_names(phj_ctx *ctx, char *buf, uint64_t cap) {
    return guard(ctx, [&] {
        if (buf == nullptr || cap == 0)
            invalid("null buffer");
        ctx->sync();
        const std::string s = ctx->timer.names();
        const uint64_t n = std::min<uint64_t>(s.size(), cap - 1);
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    });
}

phj_status phj_copy_bench(phj_ctx *ctx, void *dst, const void *src, uint64_t bytes, int reps, double *avg_ms) {
    return guard(ctx, [&] {
        if (reps <= 0)
            invalid("reps must be positive");
        launch_copy(ctx->stream, dst, src, bytes, nullptr);
        EventPair ev;
        HIPCHK(hipEventRecord(ev.a, ctx->stream));
        for (int r = 0; r < reps; ++r)
            launch_copy(ctx->stream, dst, src, bytes, ctx->tm());
        HIPCHK(hipEventRecord(ev.b, ctx->stream));
        ctx->sync();
        if (avg_ms)
            *avg_ms = ev.ms() / reps;
    });
}

phj_status phj_stream_bench(phj_ctx *ctx, void *dst, const void *src, uint64_t bytes, uint32_t mode, int reps,
                            double *avg_ms) {
    return guard(ctx, [&] {
        if (reps <= 0 || mode > PHJ_STREAM_WRITE)
            invalid("reps must be positive and mode one of PHJ_STREAM_*");
        if ((mode != PHJ_STREAM_WRITE && src == nullptr) || (mode != PHJ_STREAM_READ && dst == nullptr))
            invalid("null buffer");
        launch_stream(ctx->stream, static_cast<int>(mode), dst, src, bytes, nullptr);
        EventPair ev;
        HIPCHK(hipEventRecord(ev.a, ctx->stream));
        for (int r = 0; r < reps; ++r)
            launch_stream(ctx->stream, static_cast<int>(mode), dst, src, bytes, ctx->tm());
        HIPCHK(hipEventRecord(ev.b, ctx->stream));
        ctx->sync();
        if (avg_ms)
            *avg_ms = ev.ms() / reps;
    });
}

phj_status phj_layout_probe(phj_ctx *ctx, uint32_t mode, uint64_t bytes, uint64_t param, int reps, double *avg_ms) {
    return guard(ctx, [&] {
        if (reps <= 0 || mode > PHJ_PROBE_PASS_ROWS)
            invalid("reps must be positive and mode one of PHJ_PROBE_*");
        Scratch scr(ctx);
        std::function<void()> once;
        if (mode == PHJ_PROBE_RANDOM_READ) {
            if (bytes < 4096)
                invalid("random-read table below 4 KiB");
            const uint64_t words = 1ull << (63 - __builtin_clzll(bytes / 16));
            const uint64_t reads = param ? param : (1ull << 28);
            unsigned char *t = scr.get<unsigned char>(words * 16);
            launch_fill_u64(ctx->stream, reinterpret_cast<uint64_t *>(t), words * 2, 0x0123456789ABCDEFull);
            uint32_t *sink = scr.get<uint32_t>(4);
            uint64_t seed = 0;
            once = [=, &seed] { launch_probe_random_read(ctx->stream, t, words, reads, seed, sink); seed += reads; };
        } else if (mode == PHJ_PROBE_SCRATCH_COPY) {
            const uint64_t n = bytes & ~uint64_t{15};
            const uint64_t chunk = param & ~uint64_t{15};
            if (n == 0 || chunk == 0)
                invalid("scratch copy needs bytes and a scratch size of at least 16");
            unsigned char *src = scr.get<unsigned char>(n);
            unsigned char *dst = scr.get<unsigned char>(n);
            unsigned char *tmp = scr.get<unsigned char>(std::min(chunk, n));
            launch_fill_u64(ctx->stream, reinterpret_cast<uint64_t *>(src), n / 8, 0x0123456789ABCDEFull);
            once = [=] {
                for (uint64_t off = 0; off < n; off += chunk) {
                    const uint64_t len = std::min(chunk, n - off);
                    launch_probe_copy(ctx->stream, tmp, src + off, len);
                    launch_probe_copy(ctx->stream, dst + off, tmp, len);
                }
            };
        } else {
            const uint64_t n = bytes / 32 / 4096 * 4096;
            if (n == 0 || param == 0 || param > 4096 || (param & (param - 1)) != 0)
                invalid("pass probe needs >= 4,096 rows of 32 bytes and a power-of-two partition count <= 4,096");
            unsigned char *in = scr.get<unsigned char>(n * 32);
            unsigned char *out = scr.get<unsigned char>(n * 32);
            launch_fill_u64(ctx->stream, reinterpret_cast<uint64_t *>(in), n * 4, 0x0123456789ABCDEFull);
            const bool rows = mode == PHJ_PROBE_PASS_ROWS;
            const uint32_t parts = static_cast<uint32_t>(param);
            once = [=] { launch_probe_pass(ctx->stream, rows, in, out, n, parts); };
        }
        once(); // untimed: first touch of the buffers
        EventPair ev;
        HIPCHK(hipEventRecord(ev.a, ctx->stream));
        for (int r = 0; r < reps; ++r)
            once();
        HIPCHK(hipEventRecord(ev.b, ctx->stream));
        ctx->sync();
        if (avg_ms)
            *avg_ms = ev.ms() / reps;
    });
}

phj_status phj_placement_probe(phj_ctx *ctx, uint64_t rows, uint32_t sets, uint64_t gap, int reps, double *ms) {
    return guard(ctx, [&] {
        const uint64_t n = rows / 4096 * 4096;
        if (n == 0 || sets == 0 || sets > 8 || reps <= 0 || ms == nullptr)
            invalid("placement probe needs >= 4,096 rows, 1..8 sets, positive reps and an output array");
        // Fresh device allocations, outside the context's pool, released on return.
        std::vector<void *> raw;
        struct Release {
            std::vector<void *> &v;
            ~Release() {
                for (void *p : v)
                    (void)hipFree(p);
            }
        } release{raw};
        auto get = [&](uint64_t bytes) {
            void *p = nullptr;
            if (hipMalloc(&p, bytes) != hipSuccess)
                throw std::bad_alloc();
            raw.push_back(p);
            return p;
        };
        void *in = get(n * 32);
        launch_fill_u64(ctx->stream, static_cast<uint64_t *>(in), n * 4, 0x0123456789ABCDEFull);
        std::vector<std::array<void *, 3>> out(sets);
        for (uint32_t k = 0; k < sets; ++k) {
            if (gap == PHJ_PLACEMENT_SEPARATE) {
                out[k] = {get(n * 8), get(n * 8), get(n * 16)};
            } else {
                const uint64_t g = (gap + 255) & ~uint64_t{255};
                char *b = static_cast<char *>(get(n * 32 + 2 * g));
                out[k] = {b, b + n * 8 + g, b + n * 16 + 2 * g};
            }
        }
        for (uint32_t k = 0; k < sets; ++k) {
            auto pass = [&] {
                launch_probe_pass_cols(ctx->stream, in, static_cast<uint64_t *>(out[k][0]), static_cast<uint64_t *>(out[k][1]),
                                       out[k][2], n, 256);
            };
            pass(); // first touch
            EventPair ev;
            HIPCHK(hipEventRecord(ev.a, ctx->stream));
            for (int r = 0; r < reps; ++r)
                pass();
            HIPCHK(hipEventRecord(ev.b, ctx->stream));
            ctx->sync();
            ms[k] = ev.ms() / reps;
        }
    });
}

} // extern "C"
