// phj_types.h — the reference's data model for drop-in C++ callers.
//
// Same names, field order and widths as src/Types.h:11-108, src/Block.h:21-164
// and src/JoinOutput.h:18-84 (std::vector instead of std::pmr; no allocator
// plumbing), so code written against runCHJ/runPHJ/runPhjBep compiles against
// this header and the GPU adapter in phj_schemes.h.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace phj {

using BlockNo = uint32_t;
using RowNo = uint32_t;
inline constexpr BlockNo INVALID_BLOCK = ~BlockNo{0};

struct UInt128 {
    uint64_t low;
    uint64_t high;
    friend bool operator==(const UInt128 &a, const UInt128 &b) noexcept { return a.low == b.low && a.high == b.high; }
};
static_assert(sizeof(UInt128) == 16, "UInt128 is 16 bytes");

enum class PayloadType : uint8_t { UInt8 = 0, UInt16 = 1, UInt32 = 2, UInt64 = 3, UInt128 = 4 };

inline constexpr size_t payloadTypeSize(PayloadType t) noexcept {
    switch (t) {
    case PayloadType::UInt8: return 1;
    case PayloadType::UInt16: return 2;
    case PayloadType::UInt32: return 4;
    case PayloadType::UInt64: return 8;
    case PayloadType::UInt128: return 16;
    }
    return 0;
}

struct PayloadSchema {
    std::vector<PayloadType> types;
    size_t rowByteSize() const noexcept {
        size_t s = 0;
        for (auto t : types)
            s += payloadTypeSize(t);
        return s;
    }
};

struct RowRefCell {
    BlockNo block_no = INVALID_BLOCK;
    RowNo row_no = 0;
    bool valid() const noexcept { return block_no != INVALID_BLOCK; }
};
static_assert(sizeof(RowRefCell) == 8, "RowRefCell is 8 bytes");

inline constexpr RowRefCell INVALID_REF{INVALID_BLOCK, 0};
inline constexpr size_t PIPELINE_BLOCK_ROWS = 10'000;

struct Column {
    PayloadType type = PayloadType::UInt8;
    std::vector<std::byte> data;
    size_t elementSize() const noexcept { return payloadTypeSize(type); }
    std::byte *raw() noexcept { return data.data(); }
    const std::byte *raw() const noexcept { return data.data(); }
};

struct Block {
    size_t rows = 0;
    std::vector<uint64_t> keys;
    std::vector<Column> payloads;
    void init(const PayloadSchema &schema, size_t row_count) {
        rows = row_count;
        keys.assign(row_count, 0);
        payloads.assign(schema.types.size(), {});
        for (size_t c = 0; c < schema.types.size(); ++c) {
            payloads[c].type = schema.types[c];
            payloads[c].data.assign(row_count * payloadTypeSize(schema.types[c]), std::byte{});
        }
    }
};

struct BlockStream {
    PayloadSchema schema;
    std::vector<Block> blocks;
    size_t total_rows = 0;
};

struct OutputBlock {
    size_t rows = 0;
    std::vector<uint64_t> keys;
    std::vector<Column> left;
    std::vector<Column> right;
};

struct OutputWorker {
    std::vector<OutputBlock> blocks;
    size_t totalRows() const noexcept {
        size_t s = 0;
        for (const auto &b : blocks)
            s += b.rows;
        return s;
    }
};

struct JoinOutput {
    PayloadSchema left_schema;
    PayloadSchema right_schema;
    std::vector<OutputWorker> workers;
    size_t totalRows() const noexcept {
        size_t s = 0;
        for (const auto &w : workers)
            s += w.totalRows();
        return s;
    }
};

struct PhaseTiming {
    double wall_ms = 0.0;
    double ns_per_row = 0.0;
};

struct RadixConfig {
    std::vector<uint8_t> pass_bits;
    uint8_t totalBits() const noexcept {
        uint8_t s = 0;
        for (auto b : pass_bits)
            s = static_cast<uint8_t>(s + b);
        return s;
    }
    size_t partitions() const noexcept { return size_t{1} << totalBits(); }
};

} // namespace phj
