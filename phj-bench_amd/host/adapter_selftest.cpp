// adapter_selftest — drives runCHJ / runPHJ / runPhjBep through the C++ adapter
// the way the reference's driver does (src/Main.cpp:64-109): BlockStreams of
// PIPELINE_BLOCK_ROWS blocks in, JoinOutput out. The output is serialised as
// key | build cols | probe cols per row (src/Reference.cpp:79-110), sorted, and
// written for the caller to compare against its oracle.
//
// usage: adapter_selftest <input.bin> <output.bin>
//   input:  u64 magic, u32 scheme, u32 n_passes, u8 pass_bits[8], u64 budget_mib,
//           u64 threads, then build and probe sides: u64 rows, u32 n_cols,
//           u32 types[n_cols], u64 keys[rows], column bytes in order
//   output: u64 rows, u64 row_size, rows * row_size bytes (sorted)
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "phj_schemes.h"

using namespace phj;

namespace {

template <class T> T rd(std::ifstream &in) {
    T v{};
    in.read(reinterpret_cast<char *>(&v), sizeof(T));
    if (!in)
        throw std::runtime_error("truncated input");
    return v;
}

BlockStream readSide(std::ifstream &in) {
    BlockStream s;
    const uint64_t rows = rd<uint64_t>(in);
    const uint32_t ncols = rd<uint32_t>(in);
    for (uint32_t c = 0; c < ncols; ++c)
        s.schema.types.push_back(static_cast<PayloadType>(rd<uint32_t>(in)));
    std::vector<uint64_t> keys(rows);
    in.read(reinterpret_cast<char *>(keys.data()), static_cast<std::streamsize>(rows * 8));
    std::vector<std::vector<std::byte>> cols(ncols);
    for (uint32_t c = 0; c < ncols; ++c) {
        cols[c].resize(rows * payloadTypeSize(s.schema.types[c]));
        in.read(reinterpret_cast<char *>(cols[c].data()), static_cast<std::streamsize>(cols[c].size()));
    }
    if (!in)
        throw std::runtime_error("truncated input");
    s.total_rows = rows;
    for (uint64_t a = 0; a < rows; a += PIPELINE_BLOCK_ROWS) {
        const uint64_t n = std::min<uint64_t>(PIPELINE_BLOCK_ROWS, rows - a);
        Block b;
        b.init(s.schema, n);
        std::memcpy(b.keys.data(), keys.data() + a, n * 8);
        for (uint32_t c = 0; c < ncols; ++c) {
            const size_t sz = payloadTypeSize(s.schema.types[c]);
            std::memcpy(b.payloads[c].raw(), cols[c].data() + a * sz, n * sz);
        }
        s.blocks.push_back(std::move(b));
    }
    return s;
}

std::vector<std::string> serialize(const JoinOutput &out) {
    std::vector<std::string> rows;
    for (const auto &w : out.workers)
        for (const auto &b : w.blocks)
            for (size_t i = 0; i < b.rows; ++i) {
                std::string r(reinterpret_cast<const char *>(&b.keys[i]), 8);
                for (const auto &c : b.left)
                    r.append(reinterpret_cast<const char *>(c.raw() + i * c.elementSize()), c.elementSize());
                for (const auto &c : b.right)
                    r.append(reinterpret_cast<const char *>(c.raw() + i * c.elementSize()), c.elementSize());
                rows.push_back(std::move(r));
            }
    std::sort(rows.begin(), rows.end());
    return rows;
}

} // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <input.bin> <output.bin>\n", argv[0]);
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        if (rd<uint64_t>(in) != 0x31464c45534a4850ull) // "PHJSELF1"
            throw std::runtime_error("bad magic");
        const uint32_t scheme = rd<uint32_t>(in);
        const uint32_t npass = rd<uint32_t>(in);
        uint8_t pb[8];
        in.read(reinterpret_cast<char *>(pb), 8);
        const uint64_t budget = rd<uint64_t>(in);
        const uint64_t threads = rd<uint64_t>(in);
        const BlockStream build = readSide(in);
        const BlockStream probe = readSide(in);
        RadixConfig cfg;
        cfg.pass_bits.assign(pb, pb + npass);
        JoinOutput output;
        double e2e = 0;
        if (scheme == 1) {
            ChjResult r = runCHJ(build, probe, threads);
            e2e = r.e2e_wall_ms;
            output = std::move(r.output);
        } else if (scheme == 3) {
            PhjResult r = runPHJ(build, probe, cfg, threads);
            e2e = r.e2e_wall_ms;
            output = std::move(r.output);
        } else {
            PhjBepResult r = runPhjBep(build, probe, cfg, threads, budget);
            e2e = r.e2e_wall_ms;
            output = std::move(r.output);
        }
        const std::vector<std::string> rows = serialize(output);
        const uint64_t row_size = 8 + build.schema.rowByteSize() + probe.schema.rowByteSize();
        std::ofstream out(argv[2], std::ios::binary);
        const uint64_t n = rows.size();
        out.write(reinterpret_cast<const char *>(&n), 8);
        out.write(reinterpret_cast<const char *>(&row_size), 8);
        for (const auto &r : rows)
            out.write(r.data(), static_cast<std::streamsize>(r.size()));
        std::printf("{\"out_rows\": %llu, \"e2e_wall_ms\": %.4f, \"blocks\": %zu}\n", static_cast<unsigned long long>(n), e2e,
                    output.workers.empty() ? size_t{0} : output.workers[0].blocks.size());
        return 0;
    } catch (const std::invalid_argument &e) {
        std::fprintf(stderr, "invalid_argument: %s\n", e.what());
        return 3;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
