// ref_shim.cpp — C entry points over the compiled reference (phj-bench).
//
// TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles this file together with
// the reference's own sources where they lie (src/CHJ.cpp, src/PHJ.cpp,
// src/PHJBep.cpp, src/RadixPartition.cpp, src/Reference.cpp, src/DataGen.cpp)
// into oracle/_ref/libphjref.so. It converts flat columns into the reference's
// BlockStream (10,000-row blocks, src/Types.h:108), runs the reference code
// unchanged, and reports results through the oracle's checksum. Used to pin the
// oracle and as the CPU baseline timed beside the GPU.
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "CHJ.h"
#include "DataGen.h"
#include "Hash.h"
#include "PHJ.h"
#include "PHJBep.h"
#include "RadixPartition.h"
#include "Reference.h"
#include "oracle.h"

namespace {

phj::BlockStream toStream(const orc_relation &rel) {
    phj::BlockStream s;
    for (uint32_t c = 0; c < rel.n_cols; ++c)
        s.schema.types.push_back(static_cast<phj::PayloadType>(rel.types[c]));
    s.total_rows = rel.rows;
    const size_t br = phj::PIPELINE_BLOCK_ROWS;
    const size_t nb = (rel.rows + br - 1) / br;
    s.blocks.resize(nb);
    for (size_t b = 0; b < nb; ++b) {
        const size_t start = b * br;
        const size_t rows = std::min<size_t>(br, rel.rows - start);
        phj::Block &blk = s.blocks[b];
        blk.init(s.schema, rows);
        std::memcpy(blk.keys.data(), rel.keys + start, rows * 8);
        for (uint32_t c = 0; c < rel.n_cols; ++c) {
            const size_t sz = phj::payloadTypeSize(s.schema.types[c]);
            std::memcpy(blk.payloads[c].raw(), static_cast<const uint8_t *>(rel.cols[c]) + start * sz, rows * sz);
        }
    }
    return s;
}

void digestOutput(const phj::JoinOutput &out, orc_checksum *cs) {
    orc_checksum s{0, 0, 0};
    for (const std::string &row : phj::serializeOutput(out)) {
        uint64_t a, c;
        orc_row_digest(reinterpret_cast<const uint8_t *>(row.data()), row.size(), &a, &c);
        s.rows += 1;
        s.lane0 += a;
        s.lane1 += c;
    }
    *cs = s;
}

} // namespace

extern "C" {

typedef struct ref_result {
    uint64_t out_rows;
    orc_checksum checksum;
    double e2e_wall_ms;
    double build_shuffle_ms;
    double build_ms;
    double probe_shuffle_ms;
    double probe_ms;
    uint64_t peak_mem_bytes;
} ref_result;

void ref_int_hash64_batch(const uint64_t *in, uint64_t n, uint64_t *out) { phj::intHash64Batch(in, n, out); }

// scheme: 1 CHJ, 2 PHJ (best-effort), 3 PHJ-PURE (src/CLI.h:15-22).
int ref_run(int scheme, const orc_relation *build, const orc_relation *probe, const uint8_t *pass_bits, uint32_t n_passes,
            uint64_t threads, uint64_t bep_budget_mib, int want_checksum, ref_result *res) {
    try {
        const phj::BlockStream b = toStream(*build);
        const phj::BlockStream p = toStream(*probe);
        phj::RadixConfig cfg;
        cfg.pass_bits.assign(pass_bits, pass_bits + n_passes);
        std::memset(res, 0, sizeof(*res));
        if (scheme == 1) {
            phj::ChjResult r = phj::runCHJ(b, p, threads);
            res->out_rows = r.output.totalRows();
            res->e2e_wall_ms = r.e2e_wall_ms;
            res->build_ms = r.build.wall_ms;
            res->probe_ms = r.probe.wall_ms;
            res->peak_mem_bytes = r.peak_mem_bytes;
            if (want_checksum)
                digestOutput(r.output, &res->checksum);
        } else if (scheme == 3) {
            phj::PhjResult r = phj::runPHJ(b, p, cfg, threads);
            res->out_rows = r.output.totalRows();
            res->e2e_wall_ms = r.e2e_wall_ms;
            res->build_shuffle_ms = r.build_shuffle.wall_ms;
            res->build_ms = r.build.wall_ms;
            res->probe_shuffle_ms = r.probe_shuffle.wall_ms;
            res->probe_ms = r.probe.wall_ms;
            res->peak_mem_bytes = r.peak_mem_bytes;
            if (want_checksum)
                digestOutput(r.output, &res->checksum);
        } else if (scheme == 2) {
            phj::PhjBepResult r = phj::runPhjBep(b, p, cfg, threads, bep_budget_mib);
            res->out_rows = r.output.totalRows();
            res->e2e_wall_ms = r.e2e_wall_ms;
            res->build_shuffle_ms = r.build_shuffle.wall_ms;
            res->build_ms = r.build.wall_ms;
            res->probe_shuffle_ms = r.probe_shuffle.wall_ms;
            res->probe_ms = r.probe.wall_ms;
            res->peak_mem_bytes = r.peak_mem_bytes;
            if (want_checksum)
                digestOutput(r.output, &res->checksum);
        } else {
            return 1;
        }
    } catch (const std::exception &) {
        return 2;
    }
    return 0;
}

// radixShuffle(input, out_schema, in_col_index, cfg, threads)
// (src/RadixPartition.h:187-193) with its partitions flattened in partition
// order: out_keys/out_cols receive every partition's chain blocks in chain
// order and part_rows[p] its row count (2^total entries). Returns 0, or 2 when
// the reference throws (std::invalid_argument).
int ref_radix_shuffle(const orc_relation *in, const uint32_t *out_types, const uint64_t *in_col_index, uint32_t n_out,
                      const uint8_t *pass_bits, uint32_t n_passes, uint64_t threads, uint64_t *out_keys,
                      void *const *out_cols, uint64_t *part_rows) {
    try {
        const phj::BlockStream s = toStream(*in);
        phj::PayloadSchema schema;
        std::vector<size_t> index;
        for (uint32_t c = 0; c < n_out; ++c) {
            schema.types.push_back(static_cast<phj::PayloadType>(out_types[c]));
            index.push_back(in_col_index[c]);
        }
        phj::RadixConfig cfg;
        cfg.pass_bits.assign(pass_bits, pass_bits + n_passes);
        const phj::PartitionedShuffleOutput out = phj::radixShuffle(s, schema, index, cfg, threads);
        size_t off = 0;
        for (size_t p = 0; p < out.partitions; ++p) {
            part_rows[p] = out.partition_rows[p];
            for (const auto &blk : out.chains[p].blocks) {
                std::memcpy(out_keys + off, blk.keys.data(), blk.rows * 8);
                for (uint32_t c = 0; c < n_out; ++c) {
                    const size_t sz = phj::payloadTypeSize(schema.types[c]);
                    std::memcpy(static_cast<uint8_t *>(out_cols[c]) + off * sz, blk.payloads[c].raw(), blk.rows * sz);
                }
                off += blk.rows;
            }
        }
    } catch (const std::exception &) {
        return 2;
    }
    return 0;
}

// The reference's own oracle: referenceOutput (src/Reference.cpp:48-76).
int ref_reference_checksum(const orc_relation *build, const orc_relation *probe, orc_checksum *out) {
    const phj::BlockStream b = toStream(*build);
    const phj::BlockStream p = toStream(*probe);
    orc_checksum s{0, 0, 0};
    for (const std::string &row : phj::referenceOutput(b, p)) {
        uint64_t a, c;
        orc_row_digest(reinterpret_cast<const uint8_t *>(row.data()), row.size(), &a, &c);
        s.rows += 1;
        s.lane0 += a;
        s.lane1 += c;
    }
    *out = s;
    return 0;
}

// Sorted serialised rows of referenceOutput; returns the row count (rows are
// written only when count * row_size <= cap_bytes).
uint64_t ref_reference_rows(const orc_relation *build, const orc_relation *probe, uint8_t *buf, uint64_t cap_bytes) {
    const phj::BlockStream b = toStream(*build);
    const phj::BlockStream p = toStream(*probe);
    std::vector<std::string> rows = phj::referenceOutput(b, p);
    std::sort(rows.begin(), rows.end());
    uint64_t off = 0;
    for (const auto &r : rows) {
        if (off + r.size() <= cap_bytes)
            std::memcpy(buf + off, r.data(), r.size());
        off += r.size();
    }
    return rows.size();
}

// The reference's DataGen streams (src/DataGen.cpp:144-180), flattened.
int ref_generate(uint64_t build_rows, uint64_t probe_rows, const uint32_t *build_types, uint32_t nb_cols,
                 const uint32_t *probe_types, uint32_t np_cols, uint64_t seed, uint64_t *bkeys, void *const *bcols,
                 uint64_t *pkeys, void *const *pcols) {
    phj::PayloadSchema bs, ps;
    for (uint32_t c = 0; c < nb_cols; ++c)
        bs.types.push_back(static_cast<phj::PayloadType>(build_types[c]));
    for (uint32_t c = 0; c < np_cols; ++c)
        ps.types.push_back(static_cast<phj::PayloadType>(probe_types[c]));
    const phj::BlockStream b = phj::generateBuild(build_rows, bs, seed, 1);
    const phj::BlockStream p = phj::generateProbe(probe_rows, ps, seed, b, 1);
    auto flatten = [](const phj::BlockStream &s, uint64_t *keys, void *const *cols) {
        size_t off = 0;
        for (const auto &blk : s.blocks) {
            std::memcpy(keys + off, blk.keys.data(), blk.rows * 8);
            for (size_t c = 0; c < blk.payloads.size(); ++c) {
                const size_t sz = blk.payloads[c].elementSize();
                std::memcpy(static_cast<uint8_t *>(cols[c]) + off * sz, blk.payloads[c].raw(), blk.rows * sz);
            }
            off += blk.rows;
        }
    };
    flatten(b, bkeys, bcols);
    flatten(p, pkeys, pcols);
    return 0;
}

} // extern "C"
