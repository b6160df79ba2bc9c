// phj_schemes.cpp — runCHJ / runPHJ / runPhjBep over the C ABI (phj_gpu.h).
#include "phj_schemes.h"

#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>

#include "phj_gpu.h"

namespace phj {

namespace {

// One device context per process, created on first use.
phj_ctx *context() {
    static std::once_flag once;
    static phj_ctx *ctx = nullptr;
    static phj_status rc = PHJ_OK;
    std::call_once(once, [] { rc = phj_ctx_create(0, &ctx); });
    if (rc != PHJ_OK)
        throw std::runtime_error("no usable HIP device for the hash-join library");
    return ctx;
}

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

// Device output -> OutputBlocks of PIPELINE_BLOCK_ROWS rows.
JoinOutput fetch(phj_ctx *ctx, phj_join_result &r, const BlockStream &build, const BlockStream &probe) {
    JoinOutput out;
    out.left_schema = build.schema;
    out.right_schema = probe.schema;
    out.workers.resize(1);
    const uint64_t n = r.out_rows;
    std::vector<uint64_t> keys(n);
    std::vector<std::vector<std::byte>> left(r.n_left), right(r.n_right);
    void *lp[PHJ_MAX_COLS] = {}, *rp[PHJ_MAX_COLS] = {};
    for (uint32_t c = 0; c < r.n_left; ++c) {
        left[c].resize(n * payloadTypeSize(static_cast<PayloadType>(r.left_types[c])));
        lp[c] = left[c].data();
    }
    for (uint32_t c = 0; c < r.n_right; ++c) {
        right[c].resize(n * payloadTypeSize(static_cast<PayloadType>(r.right_types[c])));
        rp[c] = right[c].data();
    }
    check(ctx, phj_result_copy_to_host(ctx, &r, keys.data(), lp, rp));
    for (uint64_t a = 0; a < n; a += PIPELINE_BLOCK_ROWS) {
        const uint64_t rows = std::min<uint64_t>(PIPELINE_BLOCK_ROWS, n - a);
        OutputBlock ob;
        ob.rows = rows;
        ob.keys.assign(keys.begin() + a, keys.begin() + a + rows);
        auto slice = [&](std::vector<std::vector<std::byte>> &src, const PayloadSchema &sch, std::vector<Column> &dst) {
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
    check(ctx, phj_result_release(ctx, &r));
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
    phj_ctx *ctx = context();
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
    phj_ctx *ctx = context();
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
    phj_ctx *ctx = context();
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
