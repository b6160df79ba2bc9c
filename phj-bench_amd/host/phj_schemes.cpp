// phj_schemes.cpp — runCHJ / runPHJ / runPhjBep over the C ABI (phj_gpu.h).
#include "phj_schemes.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>

#include "phj_gpu.h"

namespace phj {

namespace {

// Device contexts for concurrent callers. The reference's scheme functions
// keep no shared state, so independent calls may run at once from several
// threads (src/MemTracker.h:95,168 is their only thread_local). A context
// (one HIP stream) serves one call at a time, so every call takes a free
// context from this pool, creating one when all are busy, and returns it
// when done. Contexts live until the process ends.
class ContextPool {
public:
    phj_ctx *acquire() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            if (!free_.empty()) {
                phj_ctx *c = free_.back();
                free_.pop_back();
                return c;
            }
        }
        phj_ctx *c = nullptr;
        if (phj_ctx_create(0, &c) != PHJ_OK)
            throw std::runtime_error("no usable HIP device for the hash-join library");
        return c;
    }
    void release(phj_ctx *c) {
        std::lock_guard<std::mutex> lock(mu_);
        free_.push_back(c);
    }

private:
    std::mutex mu_;
    std::vector<phj_ctx *> free_;
};

ContextPool &pool() {
    static ContextPool *p = new ContextPool; // never destroyed: contexts outlive static destructors
    return *p;
}

// The context of one call, back to the pool when the call ends (also on throw).
struct Lease {
    phj_ctx *ctx;
    Lease() : ctx(pool().acquire()) {}
    ~Lease() { pool().release(ctx); }
    Lease(const Lease &) = delete;
    Lease &operator=(const Lease &) = delete;
};

void check(phj_ctx *ctx, phj_status rc) {
    if (rc == PHJ_OK)
        return;
    const std::string msg = phj_last_error(ctx);
    if (rc == PHJ_ERR_INVALID_ARGUMENT)
        throw std::invalid_argument(msg);
    if (rc == PHJ_ERR_OUT_OF_MEMORY)
        throw std::bad_alloc();
    throw std::runtime_error(msg);
}

// A BlockStream flattened into contiguous host columns.
struct Flat {
    std::vector<uint64_t> keys;
    std::vector<std::vector<std::byte>> cols;
    phj_relation rel{};
};

void flatten(const BlockStream &s, Flat &f) {
    f.keys.resize(s.total_rows);
    f.cols.assign(s.schema.types.size(), {});
    for (size_t c = 0; c < s.schema.types.size(); ++c)
        f.cols[c].resize(s.total_rows * payloadTypeSize(s.schema.types[c]));
    size_t off = 0;
    for (const Block &b : s.blocks) {
        if (b.rows == 0)
            continue;
        if (off + b.rows > s.total_rows || b.payloads.size() != s.schema.types.size())
            throw std::invalid_argument("BlockStream rows or schema inconsistent with its blocks");
        std::memcpy(f.keys.data() + off, b.keys.data(), b.rows * 8);
        for (size_t c = 0; c < b.payloads.size(); ++c) {
            const size_t sz = payloadTypeSize(s.schema.types[c]);
            std::memcpy(f.cols[c].data() + off * sz, b.payloads[c].raw(), b.rows * sz);
        }
        off += b.rows;
    }
    if (off != s.total_rows)
        throw std::invalid_argument("BlockStream total_rows does not match its blocks");
    if (s.schema.types.size() > PHJ_MAX_COLS)
        throw std::invalid_argument("at most 8 payload columns");
    f.rel.rows = s.total_rows;
    f.rel.keys = f.keys.data();
    f.rel.n_cols = static_cast<uint32_t>(s.schema.types.size());
    f.rel.on_device = 0;
    for (size_t c = 0; c < s.schema.types.size(); ++c) {
        f.rel.cols[c].type = static_cast<uint32_t>(s.schema.types[c]);
        f.rel.cols[c].data = f.cols[c].data();
    }
}

// Device output -> OutputBlocks of PIPELINE_BLOCK_ROWS rows. The columns come
// to the host through a staging buffer of kStageBlocks blocks at a time, so
// the host holds the output once (in its blocks) plus one bounded stage.
constexpr uint64_t kStageBlocks = 64;

JoinOutput fetch(phj_ctx *ctx, phj_join_result &r, const BlockStream &build, const BlockStream &probe) {
    // the device result goes back to the context on every path out
    struct Release {
        phj_ctx *ctx;
        phj_join_result *r;
        ~Release() { (void)phj_result_release(ctx, r); }
    } release{ctx, &r};
    JoinOutput out;
    out.left_schema = build.schema;
    out.right_schema = probe.schema;
    out.workers.resize(1);
    const uint64_t n = r.out_rows;
    const uint64_t stage_rows = std::min<uint64_t>(n, kStageBlocks * PIPELINE_BLOCK_ROWS);
    std::vector<uint64_t> keys(stage_rows);
    std::vector<std::vector<std::byte>> left(r.n_left), right(r.n_right);
    void *lp[PHJ_MAX_COLS] = {}, *rp[PHJ_MAX_COLS] = {};
    for (uint32_t c = 0; c < r.n_left; ++c) {
        left[c].resize(stage_rows * payloadTypeSize(static_cast<PayloadType>(r.left_types[c])));
        lp[c] = left[c].data();
    }
    for (uint32_t c = 0; c < r.n_right; ++c) {
        right[c].resize(stage_rows * payloadTypeSize(static_cast<PayloadType>(r.right_types[c])));
        rp[c] = right[c].data();
    }
    out.workers[0].blocks.reserve((n + PIPELINE_BLOCK_ROWS - 1) / PIPELINE_BLOCK_ROWS);
    for (uint64_t s0 = 0; s0 < n; s0 += stage_rows) {
        // rows [s0, s0 + m) of every column: a view of the result shifted by s0
        const uint64_t m = std::min(stage_rows, n - s0);
        phj_join_result part = r;
        part.out_rows = m;
        part.keys = r.keys + s0;
        for (uint32_t c = 0; c < r.n_left; ++c)
            part.left[c] = static_cast<std::byte *>(r.left[c]) +
                           s0 * payloadTypeSize(static_cast<PayloadType>(r.left_types[c]));
        for (uint32_t c = 0; c < r.n_right; ++c)
            part.right[c] = static_cast<std::byte *>(r.right[c]) +
                            s0 * payloadTypeSize(static_cast<PayloadType>(r.right_types[c]));
        check(ctx, phj_result_copy_to_host(ctx, &part, keys.data(), lp, rp));
        for (uint64_t a = 0; a < m; a += PIPELINE_BLOCK_ROWS) {
            const uint64_t rows = std::min<uint64_t>(PIPELINE_BLOCK_ROWS, m - a);
            OutputBlock ob;
            ob.rows = rows;
            ob.keys.assign(keys.begin() + a, keys.begin() + a + rows);
            auto slice = [&](std::vector<std::vector<std::byte>> &src, const PayloadSchema &sch,
                             std::vector<Column> &dst) {
                dst.resize(sch.types.size());
                for (size_t c = 0; c < sch.types.size(); ++c) {
                    const size_t sz = payloadTypeSize(sch.types[c]);
                    dst[c].type = sch.types[c];
                    dst[c].data.assign(src[c].begin() + a * sz, src[c].begin() + (a + rows) * sz);
                }
            };
            slice(left, build.schema, ob.left);
            slice(right, probe.schema, ob.right);
            out.workers[0].blocks.push_back(std::move(ob));
        }
    }
    return out;
}

PhaseTiming phase(const phj_phase_timing &p) { return PhaseTiming{p.wall_ms, p.ns_per_row}; }

phj_join_result join(phj_ctx *ctx, uint32_t scheme, const BlockStream &build, const BlockStream &probe,
                     const RadixConfig *cfg, size_t threads, size_t budget_mib, Flat &fb, Flat &fp) {
    flatten(build, fb);
    flatten(probe, fp);
    phj_join_desc d{};
    d.scheme = scheme;
    d.threads = static_cast<uint32_t>(threads == 0 ? 1 : threads);
    d.build = fb.rel;
    d.probe = fp.rel;
    if (cfg != nullptr) {
        if (cfg->pass_bits.size() > PHJ_MAX_PASSES)
            throw std::invalid_argument("at most 8 passes");
        d.n_passes = static_cast<uint32_t>(cfg->pass_bits.size());
        for (size_t i = 0; i < cfg->pass_bits.size(); ++i)
            d.pass_bits[i] = cfg->pass_bits[i];
    }
    d.bep_budget_bytes = static_cast<uint64_t>(budget_mib) << 20;
    phj_join_result r{};
    check(ctx, phj_join(ctx, &d, &r));
    return r;
}

} // namespace

ChjResult runCHJ(const BlockStream &build, const BlockStream &probe, size_t threads) {
    const Lease lease;
    phj_ctx *ctx = lease.ctx;
    Flat fb, fp;
    phj_join_result r = join(ctx, PHJ_SCHEME_CHJ, build, probe, nullptr, threads, 0, fb, fp);
    ChjResult res;
    res.build = phase(r.build);
    res.probe = phase(r.probe);
    res.e2e_wall_ms = r.e2e_wall_ms;
    res.peak_mem_bytes = r.peak_mem_bytes;
    res.output = fetch(ctx, r, build, probe);
    return res;
}

PhjResult runPHJ(const BlockStream &build, const BlockStream &probe, const RadixConfig &cfg, size_t threads) {
    const Lease lease;
    phj_ctx *ctx = lease.ctx;
    Flat fb, fp;
    phj_join_result r = join(ctx, PHJ_SCHEME_PHJ_PURE, build, probe, &cfg, threads, 0, fb, fp);
    PhjResult res;
    res.build_shuffle = phase(r.build_shuffle);
    res.build = phase(r.build);
    res.probe_shuffle = phase(r.probe_shuffle);
    res.probe = phase(r.probe);
    res.e2e_wall_ms = r.e2e_wall_ms;
    res.peak_mem_bytes = r.peak_mem_bytes;
    res.output = fetch(ctx, r, build, probe);
    return res;
}

PhjBepResult runPhjBep(const BlockStream &build, const BlockStream &probe, const RadixConfig &cfg, size_t threads,
                       size_t bep_budget_mib) {
    const Lease lease;
    phj_ctx *ctx = lease.ctx;
    Flat fb, fp;
    phj_join_result r = join(ctx, PHJ_SCHEME_PHJ, build, probe, &cfg, threads, bep_budget_mib, fb, fp);
    PhjBepResult res;
    res.build_shuffle = phase(r.build_shuffle);
    res.build = phase(r.build);
    res.probe_shuffle = phase(r.probe_shuffle);
    res.probe = phase(r.probe);
    res.eviction_overhead = phase(r.eviction_overhead);
    res.e2e_wall_ms = r.e2e_wall_ms;
    res.peak_mem_bytes = r.peak_mem_bytes;
    res.bep_budget_mib = r.bep_budget_mib;
    res.bep_evictions = r.bep_evictions;
    res.bep_refinements = r.bep_refinements;
    res.bep_peak_bytes = r.bep_peak_bytes;
    res.bep_build_skip_retries = r.bep_build_skip_retries;
    res.output = fetch(ctx, r, build, probe);
    return res;
}

} // namespace phj
