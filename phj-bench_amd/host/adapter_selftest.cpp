// adapter_selftest — drives runCHJ / runPHJ / runPhjBep through the C++ adapter
// the way the reference's driver does (src/Main.cpp:64-109): BlockStreams of
// PIPELINE_BLOCK_ROWS blocks in, JoinOutput out. The output is serialised as
// key | build cols | probe cols per row (src/Reference.cpp:79-110), sorted, and
// written for the caller to compare against its oracle.
//
// usage: adapter_selftest <input.bin> <output.bin> [callers]
//   callers: run the join from this many threads at once (default 1); every
//            caller's sorted rows must equal the first caller's, which are
//            written out
//   input:  u64 magic, u32 scheme, u32 n_passes, u8 pass_bits[8], u64 budget_mib,
//           u64 threads, then build and probe sides: u64 rows, u32 n_cols,
//           u32 types[n_cols], u64 keys[rows], column bytes in order
//   output: u64 rows, u64 row_size, rows * row_size bytes (sorted)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <thread>
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

// One call of the chosen scheme, as the reference's driver makes it.
JoinOutput runScheme(uint32_t scheme, const BlockStream &build, const BlockStream &probe, const RadixConfig &cfg,
                     uint64_t threads, uint64_t budget, double &e2e) {
    if (scheme == 1) {
        ChjResult r = runCHJ(build, probe, threads);
        e2e = r.e2e_wall_ms;
        return std::move(r.output);
    }
    if (scheme == 3) {
        PhjResult r = runPHJ(build, probe, cfg, threads);
        e2e = r.e2e_wall_ms;
        return std::move(r.output);
    }
    PhjBepResult r = runPhjBep(build, probe, cfg, threads, budget);
    e2e = r.e2e_wall_ms;
    return std::move(r.output);
}

} // namespace

int main(int argc, char **argv) {
    if (argc != 3 && argc != 4) {
        std::fprintf(stderr, "usage: %s <input.bin> <output.bin> [callers]\n", argv[0]);
        return 2;
    }
    try {
        const int callers = argc == 4 ? std::max(1, std::atoi(argv[3])) : 1;
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
        // every caller starts its call together, each on its own thread
        std::vector<std::vector<std::string>> rows(callers);
        std::vector<size_t> blocks(callers, 0);
        std::vector<double> e2e(callers, 0.0);
        std::vector<std::string> errors(callers);
        std::vector<int> invalid(callers, 0);
        auto call = [&](int t) {
            try {
                const JoinOutput output = runScheme(scheme, build, probe, cfg, threads, budget, e2e[t]);
                blocks[t] = output.workers.empty() ? size_t{0} : output.workers[0].blocks.size();
                rows[t] = serialize(output);
            } catch (const std::invalid_argument &e) {
                invalid[t] = 1;
                errors[t] = e.what();
            } catch (const std::exception &e) {
                errors[t] = e.what();
            }
        };
        if (callers == 1) {
            call(0);
        } else {
            std::vector<std::thread> ts;
            for (int t = 0; t < callers; ++t)
                ts.emplace_back(call, t);
            for (auto &t : ts)
                t.join();
        }
        for (int t = 0; t < callers; ++t) {
            if (invalid[t])
                throw std::invalid_argument(errors[t]);
            if (!errors[t].empty())
                throw std::runtime_error("caller " + std::to_string(t) + ": " + errors[t]);
            if (rows[t] != rows[0])
                throw std::runtime_error("caller " + std::to_string(t) + " returned other rows than caller 0");
        }
        const uint64_t row_size = 8 + build.schema.rowByteSize() + probe.schema.rowByteSize();
        std::ofstream out(argv[2], std::ios::binary);
        const uint64_t n = rows[0].size();
        out.write(reinterpret_cast<const char *>(&n), 8);
        out.write(reinterpret_cast<const char *>(&row_size), 8);
        for (const auto &r : rows[0])
            out.write(r.data(), static_cast<std::streamsize>(r.size()));
        std::printf("{\"out_rows\": %llu, \"e2e_wall_ms\": %.4f, \"blocks\": %zu, \"callers\": %d}\n",
                    static_cast<unsigned long long>(n), e2e[0], blocks[0], callers);
        return 0;
    } catch (const std::invalid_argument &e) {
        std::fprintf(stderr, "invalid_argument: %s\n", e.what());
        return 3;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
