// oracle.cpp — CPU oracle for the phj-bench hash-join hot path.
//
// TEST INFRASTRUCTURE ONLY: this is the checker. Only tests/, the smoke check
// and the benchmark's CPU-baseline leg load it. The product library never
// depends on it. Each function names the reference lines it restates; the
// header lists how every function is pinned.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;
constexpr uint64_t kMask63 = 0x7FFFFFFFFFFFFFFFULL;

// src/Hash.h:18-26 — intHash64: salt xor, then the Murmur3 fmix64 steps.
inline uint64_t intHash64(uint64_t x) {
    x ^= 0x4cf5ad432745937fULL;
    x = (x ^ (x >> 33)) * 0xff51afd7ed558ccdULL;
    x = (x ^ (x >> 33)) * 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

// SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014).
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline uint64_t streamBase(uint64_t seed, uint64_t stream) {
    return mix64(seed ^ mix64(stream * kGolden + 0x632BE59BD9B4E019ULL));
}

inline uint64_t rand64(uint64_t base, uint64_t i) { return mix64(base + (i + 1) * kGolden); }

inline uint64_t mulhi(uint64_t a, uint64_t b) {
    return static_cast<uint64_t>((static_cast<unsigned __int128>(a) * b) >> 64);
}

// Stream ids of the harness generator (the GPU generator uses the same ids).
enum : uint64_t {
    kStreamBuildKey = 1,
    kStreamProbeSelect = 2,
    kStreamProbePick = 3,
    kStreamProbeMiss = 4,
    kStreamZipfDup = 5,
    kStreamShuffle = 6,
    kStreamZipfRank = 7,
    kStreamBuildPay = 0x100,
    kStreamBuildPayHi = 0x180,
    kStreamProbePay = 0x200,
    kStreamProbePayHi = 0x280,
};

// Keyed bijection on [0, 2^63) fixing 0: xorshifts and odd multiplies mod 2^63.
struct Perm63 {
    uint64_t m1, m2, m3;
    explicit Perm63(uint64_t base)
        : m1((rand64(base, 0) | 1) & kMask63), m2((rand64(base, 1) | 1) & kMask63), m3((rand64(base, 2) | 1) & kMask63) {}
    uint64_t operator()(uint64_t x) const {
        x &= kMask63;
        x ^= x >> 31;
        x = (x * m1) & kMask63;
        x ^= x >> 29;
        x = (x * m2) & kMask63;
        x ^= x >> 32;
        x = (x * m3) & kMask63;
        x ^= x >> 30;
        return x;
    }
};

// Keyed bijection on [0, N) by cycle walking over a 2^m domain.
struct ShufflePerm {
    uint64_t n, mask, o1, o2, c1;
    unsigned r;
    ShufflePerm(uint64_t base, uint64_t n_) : n(n_) {
        unsigned m = 0;
        while ((uint64_t{1} << m) < n && m < 63)
            ++m;
        mask = m == 0 ? 0 : ((uint64_t{1} << m) - 1);
        r = (m + 1) / 2;
        if (r == 0)
            r = 1;
        o1 = rand64(base, 10) | 1;
        o2 = rand64(base, 11) | 1;
        c1 = rand64(base, 12);
    }
    uint64_t step(uint64_t x) const {
        x = (x * o1 + c1) & mask;
        x ^= x >> r;
        x = (x * o2) & mask;
        x ^= x >> r;
        return x;
    }
    uint64_t operator()(uint64_t x) const {
        if (mask == 0)
            return 0;
        uint64_t y = step(x);
        while (y >= n)
            y = step(y);
        return y;
    }
};

// Zipf(s) CDF over ranks 1..n as 53-bit integer thresholds; rank k is drawn
// when u < t[k-1] for the first such k, u uniform on [0, 2^53).
std::vector<uint64_t> zipfThresholds(uint64_t n, double s) {
    std::vector<double> cum(n);
    double acc = 0.0;
    for (uint64_t k = 1; k <= n; ++k) {
        acc += std::pow(static_cast<double>(k), -s);
        cum[k - 1] = acc;
    }
    std::vector<uint64_t> t(n);
    const double scale = 9007199254740992.0; // 2^53
    for (uint64_t k = 0; k < n; ++k)
        t[k] = static_cast<uint64_t>(cum[k] / acc * scale);
    if (n > 0)
        t[n - 1] = uint64_t{1} << 53;
    return t;
}

inline uint64_t zipfDraw(const std::vector<uint64_t> &t, uint64_t r) {
    const uint64_t u = r >> 11;
    return static_cast<uint64_t>(std::upper_bound(t.begin(), t.end(), u) - t.begin()); // 0-based rank index
}

// src/DataGen.cpp:15-19 — salts of the reference's default streams.
constexpr uint64_t kRefBuildKeySalt = 0x6f1bda28e54e2bd1ULL;
constexpr uint64_t kRefProbePickSalt = 0xc52f01f0e0c46cd3ULL;
constexpr uint64_t kRefBuildPaySalt = 0xf3e9aa18b6517527ULL;
constexpr uint64_t kRefProbePaySalt = 0x9e3779b97f4a7c15ULL;
constexpr uint64_t kRefColumnSaltStride = 0x9ddfea08eb382d69ULL;

size_t typeSize(uint32_t t) {
    switch (t) {
    case 0: return 1;
    case 1: return 2;
    case 2: return 4;
    case 3: return 8;
    case 4: return 16;
    default: return 0;
    }
}

inline void storeValue(void *col, uint32_t type, uint64_t idx, uint64_t lo, uint64_t hi) {
    const size_t sz = typeSize(type);
    auto *dst = static_cast<uint8_t *>(col) + idx * sz;
    if (sz <= 8) {
        std::memcpy(dst, &lo, sz); // little-endian truncation, src/DataGen.cpp:24-52
    } else {
        std::memcpy(dst, &lo, 8);
        std::memcpy(dst + 8, &hi, 8);
    }
}

// Everything a generator needs for one spec, derived once.
struct GenState {
    orc_gen_spec spec;
    uint64_t b_key, b_sel, b_pick, b_miss, b_dup, b_shuf, b_rank;
    uint64_t b_bpay[ORC_MAX_COLS], b_bpayhi[ORC_MAX_COLS], b_ppay[ORC_MAX_COLS], b_ppayhi[ORC_MAX_COLS];
    Perm63 perm;
    // Zipf only
    std::vector<uint64_t> dup_t, rank_t, prefix;
    uint64_t build_total = 0;

    explicit GenState(const orc_gen_spec &s)
        : spec(s), b_key(streamBase(s.seed, kStreamBuildKey)), b_sel(streamBase(s.seed, kStreamProbeSelect)),
          b_pick(streamBase(s.seed, kStreamProbePick)), b_miss(streamBase(s.seed, kStreamProbeMiss)),
          b_dup(streamBase(s.seed, kStreamZipfDup)), b_shuf(streamBase(s.seed, kStreamShuffle)),
          b_rank(streamBase(s.seed, kStreamZipfRank)), perm(streamBase(s.seed, kStreamBuildKey)) {
        for (uint64_t c = 0; c < ORC_MAX_COLS; ++c) {
            b_bpay[c] = streamBase(s.seed, kStreamBuildPay + c);
            b_bpayhi[c] = streamBase(s.seed, kStreamBuildPayHi + c);
            b_ppay[c] = streamBase(s.seed, kStreamProbePay + c);
            b_ppayhi[c] = streamBase(s.seed, kStreamProbePayHi + c);
        }
        build_total = s.build_rows;
        if (s.distribution == 2) {
            const uint64_t maxdup = s.zipf_build_max_dup == 0 ? 1 : s.zipf_build_max_dup;
            dup_t = zipfThresholds(maxdup, s.zipf_build_s_milli / 1000.0);
            rank_t = zipfThresholds(s.build_rows, s.zipf_probe_s_milli / 1000.0);
            prefix.assign(s.build_rows + 1, 0);
            for (uint64_t d = 0; d < s.build_rows; ++d)
                prefix[d + 1] = prefix[d] + 1 + zipfDraw(dup_t, rand64(b_dup, d));
            build_total = prefix[s.build_rows];
        }
    }

    uint64_t buildKeyOfIndex(uint64_t i) const { return perm(i + 1); }

    // Build row at position i: key and the distinct-key index it carries.
    uint64_t buildKey(uint64_t i) const {
        switch (spec.distribution) {
        case 1: return intHash64((spec.seed ^ kRefBuildKeySalt) + i); // src/DataGen.cpp:86,94
        case 2: {
            const ShufflePerm sp(b_shuf, build_total);
            const uint64_t u = sp(i);
            const uint64_t d = static_cast<uint64_t>(std::upper_bound(prefix.begin(), prefix.end(), u) - prefix.begin()) - 1;
            return buildKeyOfIndex(d);
        }
        default: return buildKeyOfIndex(i);
        }
    }

    void payload(int side, uint32_t c, uint64_t i, uint64_t &lo, uint64_t &hi) const {
        const uint32_t type = side == 0 ? spec.build_types[c] : spec.probe_types[c];
        if (spec.distribution == 1) {
            // src/DataGen.cpp:95-101 and :129-135
            const uint64_t pay_salt = spec.seed ^ (side == 0 ? kRefBuildPaySalt : kRefProbePaySalt);
            const uint64_t col_salt = pay_salt ^ (kRefColumnSaltStride * (c + 1));
            lo = intHash64(col_salt ^ i);
            hi = type == 4 ? intHash64(col_salt ^ (i + (uint64_t{1} << 33))) : 0;
            return;
        }
        lo = rand64(side == 0 ? b_bpay[c] : b_ppay[c], i);
        hi = type == 4 ? rand64(side == 0 ? b_bpayhi[c] : b_ppayhi[c], i) : 0;
    }

    // Probe row j: key plus the build row it was drawn from (or ~0 on a miss).
    uint64_t probeKey(uint64_t j, uint64_t &pick) const {
        if (spec.distribution == 1) {
            // src/DataGen.cpp:117-128 — pick = intHash64(pick_salt + i) % build_rows
            pick = intHash64((spec.seed ^ kRefProbePickSalt) + j) % spec.build_rows;
            return buildKey(pick);
        }
        const bool matched = mulhi(rand64(b_sel, j), 1000000) < spec.match_ppm;
        if (!matched) {
            pick = ~uint64_t{0};
            return (uint64_t{1} << 63) | (rand64(b_miss, j) >> 1);
        }
        if (spec.distribution == 2) {
            const uint64_t d = zipfDraw(rank_t, rand64(b_rank, j));
            pick = d; // distinct-key index, not a row position
            return buildKeyOfIndex(d);
        }
        pick = mulhi(rand64(b_pick, j), spec.build_rows);
        return buildKeyOfIndex(pick);
    }
};

template <class F> void parallelFor(uint64_t n, uint32_t threads, F &&f) {
    if (threads <= 1 || n < 4096) {
        f(uint64_t{0}, n);
        return;
    }
    std::vector<std::thread> ts;
    ts.reserve(threads);
    for (uint32_t t = 0; t < threads; ++t) {
        const uint64_t a = n * t / threads, b = n * (t + 1) / threads;
        ts.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto &t : ts)
        t.join();
}

// Row digest over the serialised bytes: two chained SplitMix64 lanes over the
// little-endian 8-byte words, final word zero-padded.
struct Digest {
    uint64_t a, c;
    explicit Digest(uint64_t len) : a(0x243F6A8885A308D3ULL ^ len), c(0x13198A2E03707344ULL + len) {}
    void word(uint64_t w) {
        a = mix64(a ^ w);
        c = mix64(c + w * kGolden);
    }
};

// Streams field bytes into 8-byte words.
struct WordPacker {
    Digest &d;
    uint64_t acc = 0;
    unsigned fill = 0; // bytes in acc
    explicit WordPacker(Digest &dg) : d(dg) {}
    void bytes(const uint8_t *p, size_t n) {
        for (size_t i = 0; i < n; ++i) {
            acc |= static_cast<uint64_t>(p[i]) << (8 * fill);
            if (++fill == 8) {
                d.word(acc);
                acc = 0;
                fill = 0;
            }
        }
    }
    void finish() {
        if (fill != 0)
            d.word(acc);
        acc = 0;
        fill = 0;
    }
};

} // namespace

extern "C" {

size_t orc_type_size(uint32_t type) { return typeSize(type); }

uint64_t orc_int_hash64(uint64_t x) { return intHash64(x); }

void orc_int_hash64_batch(const uint64_t *in, uint64_t n, uint64_t *out) {
    for (uint64_t i = 0; i < n; ++i)
        out[i] = intHash64(in[i]);
}

uint64_t orc_gen_build_rows(const orc_gen_spec *spec) {
    if (spec->distribution != 2)
        return spec->build_rows;
    GenState g(*spec);
    return g.build_total;
}

int orc_gen_fill(const orc_gen_spec *spec, int side, uint64_t i0, uint64_t n, uint64_t *keys, void *const *cols,
                 uint32_t threads) {
    GenState g(*spec);
    const uint32_t ncols = side == 0 ? spec->build_n_cols : spec->probe_n_cols;
    const uint8_t *types = side == 0 ? spec->build_types : spec->probe_types;
    parallelFor(n, threads, [&](uint64_t a, uint64_t b) {
        for (uint64_t k = a; k < b; ++k) {
            const uint64_t i = i0 + k;
            if (keys != nullptr) {
                uint64_t pick;
                keys[k] = side == 0 ? g.buildKey(i) : g.probeKey(i, pick);
            }
            for (uint32_t c = 0; c < ncols; ++c) {
                if (cols == nullptr || cols[c] == nullptr)
                    continue;
                uint64_t lo, hi;
                g.payload(side, c, i, lo, hi);
                storeValue(cols[c], types[c], k, lo, hi);
            }
        }
    });
    return 0;
}

// src/Reference.cpp:48-76 — key -> [(block,row)] index in build order, then
// per probe row emit one row per indexed build row in insertion order.
uint64_t orc_reference_join_count(const uint64_t *build_keys, uint64_t nb, const uint64_t *probe_keys, uint64_t np) {
    std::unordered_map<uint64_t, uint64_t> counts;
    counts.reserve(nb * 2);
    for (uint64_t r = 0; r < nb; ++r)
        ++counts[build_keys[r]];
    uint64_t total = 0;
    for (uint64_t j = 0; j < np; ++j) {
        auto it = counts.find(probe_keys[j]);
        if (it != counts.end())
            total += it->second;
    }
    return total;
}

uint64_t orc_reference_join_pairs(const uint64_t *build_keys, uint64_t nb, const uint64_t *probe_keys, uint64_t np,
                                  uint64_t *probe_idx, uint64_t *build_idx, uint64_t cap) {
    std::unordered_map<uint64_t, std::vector<uint64_t>> index;
    index.reserve(nb * 2);
    for (uint64_t r = 0; r < nb; ++r)
        index[build_keys[r]].push_back(r);
    uint64_t m = 0;
    for (uint64_t j = 0; j < np; ++j) {
        auto it = index.find(probe_keys[j]);
        if (it == index.end())
            continue;
        for (uint64_t r : it->second) {
            if (m < cap) {
                probe_idx[m] = j;
                build_idx[m] = r;
            }
            ++m;
        }
    }
    return m;
}

uint64_t orc_row_size(const uint32_t *left_types, uint32_t n_left, const uint32_t *right_types, uint32_t n_right) {
    uint64_t s = 8;
    for (uint32_t c = 0; c < n_left; ++c)
        s += typeSize(left_types[c]);
    for (uint32_t c = 0; c < n_right; ++c)
        s += typeSize(right_types[c]);
    return s;
}

// src/Reference.cpp:22-43 and :79-110 — key (8 B, little-endian), then each
// build column, then each probe column, at its element width.
void orc_serialize_rows(uint64_t n, const uint64_t *keys, const uint32_t *left_types, const void *const *left,
                        uint32_t n_left, const uint32_t *right_types, const void *const *right, uint32_t n_right,
                        uint8_t *out) {
    const uint64_t rs = orc_row_size(left_types, n_left, right_types, n_right);
    for (uint64_t i = 0; i < n; ++i) {
        uint8_t *dst = out + i * rs;
        std::memcpy(dst, &keys[i], 8);
        dst += 8;
        for (uint32_t c = 0; c < n_left; ++c) {
            const size_t sz = typeSize(left_types[c]);
            std::memcpy(dst, static_cast<const uint8_t *>(left[c]) + i * sz, sz);
            dst += sz;
        }
        for (uint32_t c = 0; c < n_right; ++c) {
            const size_t sz = typeSize(right_types[c]);
            std::memcpy(dst, static_cast<const uint8_t *>(right[c]) + i * sz, sz);
            dst += sz;
        }
    }
}

// src/Reference.cpp:113-133 — both sides sorted as byte strings.
void orc_sort_rows(uint8_t *rows, uint64_t n, uint64_t row_size) {
    std::vector<uint64_t> idx(n);
    std::iota(idx.begin(), idx.end(), uint64_t{0});
    std::sort(idx.begin(), idx.end(),
              [&](uint64_t a, uint64_t b) { return std::memcmp(rows + a * row_size, rows + b * row_size, row_size) < 0; });
    std::vector<uint8_t> tmp(n * row_size);
    for (uint64_t i = 0; i < n; ++i)
        std::memcpy(tmp.data() + i * row_size, rows + idx[i] * row_size, row_size);
    std::memcpy(rows, tmp.data(), tmp.size());
}

void orc_row_digest(const uint8_t *row, uint64_t row_size, uint64_t *lane0, uint64_t *lane1) {
    Digest d(row_size);
    WordPacker p(d);
    p.bytes(row, row_size);
    p.finish();
    *lane0 = d.a;
    *lane1 = d.c;
}

void orc_checksum_rows(const uint8_t *rows, uint64_t n, uint64_t row_size, orc_checksum *out) {
    orc_checksum s{0, 0, 0};
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t a, c;
        orc_row_digest(rows + i * row_size, row_size, &a, &c);
        s.rows += 1;
        s.lane0 += a;
        s.lane1 += c;
    }
    *out = s;
}

void orc_checksum_columns(uint64_t n, const uint64_t *keys, const uint32_t *left_types, const void *const *left,
                          uint32_t n_left, const uint32_t *right_types, const void *const *right, uint32_t n_right,
                          orc_checksum *out) {
    const uint64_t rs = orc_row_size(left_types, n_left, right_types, n_right);
    std::vector<uint8_t> row(rs);
    orc_checksum s{0, 0, 0};
    for (uint64_t i = 0; i < n; ++i) {
        orc_serialize_rows(1, keys + i, left_types, nullptr, 0, right_types, nullptr, 0, row.data());
        uint8_t *dst = row.data() + 8;
        for (uint32_t c = 0; c < n_left; ++c) {
            const size_t sz = typeSize(left_types[c]);
            std::memcpy(dst, static_cast<const uint8_t *>(left[c]) + i * sz, sz);
            dst += sz;
        }
        for (uint32_t c = 0; c < n_right; ++c) {
            const size_t sz = typeSize(right_types[c]);
            std::memcpy(dst, static_cast<const uint8_t *>(right[c]) + i * sz, sz);
            dst += sz;
        }
        uint64_t a, cc;
        orc_row_digest(row.data(), rs, &a, &cc);
        s.rows += 1;
        s.lane0 += a;
        s.lane1 += cc;
    }
    *out = s;
}

int orc_reference_join_checksum(const orc_relation *build, const orc_relation *probe, orc_checksum *out) {
    std::unordered_map<uint64_t, std::vector<uint64_t>> index;
    index.reserve(build->rows * 2);
    for (uint64_t r = 0; r < build->rows; ++r)
        index[build->keys[r]].push_back(r);
    const uint64_t rs = orc_row_size(build->types, build->n_cols, probe->types, probe->n_cols);
    std::vector<uint8_t> row(rs);
    orc_checksum s{0, 0, 0};
    for (uint64_t j = 0; j < probe->rows; ++j) {
        auto it = index.find(probe->keys[j]);
        if (it == index.end())
            continue;
        for (uint64_t r : it->second) {
            uint8_t *dst = row.data();
            std::memcpy(dst, &probe->keys[j], 8);
            dst += 8;
            for (uint32_t c = 0; c < build->n_cols; ++c) {
                const size_t sz = typeSize(build->types[c]);
                std::memcpy(dst, static_cast<const uint8_t *>(build->cols[c]) + r * sz, sz);
                dst += sz;
            }
            for (uint32_t c = 0; c < probe->n_cols; ++c) {
                const size_t sz = typeSize(probe->types[c]);
                std::memcpy(dst, static_cast<const uint8_t *>(probe->cols[c]) + j * sz, sz);
                dst += sz;
            }
            uint64_t a, cc;
            orc_row_digest(row.data(), rs, &a, &cc);
            s.rows += 1;
            s.lane0 += a;
            s.lane1 += cc;
        }
    }
    *out = s;
    return 0;
}

int orc_gen_expected_checksum(const orc_gen_spec *spec, uint32_t threads, orc_checksum *out) {
    GenState g(*spec);
    // Zipf: build row of each shuffled slot, tabulated from the forward walk
    // (slot_row[sp(i)] = i), so a probe hit on distinct key d pairs with rows
    // slot_row[prefix[d] .. prefix[d+1]).
    std::vector<uint64_t> slot_row;
    if (spec->distribution == 2) {
        const ShufflePerm sp(g.b_shuf, g.build_total);
        slot_row.assign(g.build_total, 0);
        for (uint64_t i = 0; i < g.build_total; ++i)
            slot_row[sp(i)] = i;
    }
    uint32_t lt[ORC_MAX_COLS], rt[ORC_MAX_COLS];
    for (uint32_t c = 0; c < ORC_MAX_COLS; ++c) {
        lt[c] = spec->build_types[c];
        rt[c] = spec->probe_types[c];
    }
    const uint64_t rs = orc_row_size(lt, spec->build_n_cols, rt, spec->probe_n_cols);
    const uint32_t nt = threads == 0 ? 1 : threads;
    std::vector<orc_checksum> part(nt, orc_checksum{0, 0, 0});
    const uint64_t np = spec->probe_rows;
    auto work = [&](uint32_t t) {
        const uint64_t a = np * t / nt, b = np * (t + 1) / nt;
        std::vector<uint8_t> row(rs);
        orc_checksum s{0, 0, 0};
        for (uint64_t j = a; j < b; ++j) {
            uint64_t pick;
            const uint64_t key = g.probeKey(j, pick);
            if (pick == ~uint64_t{0} || spec->build_rows == 0)
                continue;
            uint64_t y0 = 0, y1 = 1; // build rows: pick itself, or the slots of distinct key pick
            if (spec->distribution == 2) {
                y0 = g.prefix[pick];
                y1 = g.prefix[pick + 1];
            }
            for (uint64_t y = y0; y < y1; ++y) {
                const uint64_t brow = spec->distribution == 2 ? slot_row[y] : pick;
                uint8_t *dst = row.data();
                std::memcpy(dst, &key, 8);
                dst += 8;
                for (uint32_t c = 0; c < spec->build_n_cols; ++c) {
                    uint64_t lo, hi;
                    g.payload(0, c, brow, lo, hi);
                    const size_t sz = typeSize(lt[c]);
                    uint8_t v[16];
                    std::memcpy(v, &lo, 8);
                    std::memcpy(v + 8, &hi, 8);
                    std::memcpy(dst, v, sz);
                    dst += sz;
                }
                for (uint32_t c = 0; c < spec->probe_n_cols; ++c) {
                    uint64_t lo, hi;
                    g.payload(1, c, j, lo, hi);
                    const size_t sz = typeSize(rt[c]);
                    uint8_t v[16];
                    std::memcpy(v, &lo, 8);
                    std::memcpy(v + 8, &hi, 8);
                    std::memcpy(dst, v, sz);
                    dst += sz;
                }
                uint64_t la, lc;
                orc_row_digest(row.data(), rs, &la, &lc);
                s.rows += 1;
                s.lane0 += la;
                s.lane1 += lc;
            }
        }
        part[t] = s;
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> ts;
        for (uint32_t t = 0; t < nt; ++t)
            ts.emplace_back(work, t);
        for (auto &t : ts)
            t.join();
    }
    orc_checksum s{0, 0, 0};
    for (auto &p : part) {
        s.rows += p.rows;
        s.lane0 += p.lane0;
        s.lane1 += p.lane1;
    }
    *out = s;
    return 0;
}

} // extern "C"
