// gen.h — synthetic workload generators (harness side of libphj_hip.so).
//
// Closed forms of (seed, stream, row index), so any row can be recomputed
// independently: uniform distinct build keys in [1, 2^63) via a keyed 63-bit
// bijection, probe keys drawn from the build keys at a match rate with misses
// in [2^63, 2^64), Zipf-duplicated build keys with Zipf-ranked probes, and the
// reference's own DataGen streams (src/DataGen.cpp:84-139).
#pragma once

#include <stdint.h>

#include "kernels.h"

namespace phjd {

enum GenStream : uint64_t {
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

// src/DataGen.cpp:15-19 — the reference streams' salts.
constexpr uint64_t kRefBuildKeySalt = 0x6f1bda28e54e2bd1ull;
constexpr uint64_t kRefProbePickSalt = 0xc52f01f0e0c46cd3ull;
constexpr uint64_t kRefBuildPaySalt = 0xf3e9aa18b6517527ull;
constexpr uint64_t kRefProbePaySalt = 0x9e3779b97f4a7c15ull;
constexpr uint64_t kRefColumnSaltStride = 0x9ddfea08eb382d69ull;

__host__ __device__ __forceinline__ uint64_t stream_base(uint64_t seed, uint64_t stream) {
    return mix64(seed ^ mix64(stream * kGolden + 0x632BE59BD9B4E019ull));
}

struct GenParams {
    uint64_t seed;
    uint64_t build_rows;  // distinct keys (Zipf) or rows
    uint64_t build_total; // rows on the build side
    uint64_t probe_rows;
    uint32_t dist;
    uint32_t match_ppm;
    uint32_t build_types[kMaxCols];
    uint32_t probe_types[kMaxCols];
    uint32_t n_build_cols, n_probe_cols;
    uint64_t b_sel, b_pick, b_miss, b_rank;
    uint64_t b_bpay[kMaxCols], b_bpayhi[kMaxCols], b_ppay[kMaxCols], b_ppayhi[kMaxCols];
    uint64_t m1, m2, m3; // 63-bit bijection multipliers
    // cycle-walking shuffle of the Zipf build rows (sh_i1, sh_i2: inverses of
    // sh_o1, sh_o2 modulo 2^64, for the inverse walk of the closed form)
    uint64_t sh_mask, sh_o1, sh_o2, sh_c1, sh_i1, sh_i2;
    uint32_t sh_r;
    uint32_t n_rank;
    const uint64_t *prefix; // Zipf: distinct key d owns shuffled slots [prefix[d], prefix[d+1])
    const uint64_t *rank_t; // Zipf: probe rank thresholds (53-bit)
};

__host__ __device__ __forceinline__ uint64_t perm63(const GenParams &g, uint64_t x) {
    x &= kMask63;
    x ^= x >> 31;
    x = (x * g.m1) & kMask63;
    x ^= x >> 29;
    x = (x * g.m2) & kMask63;
    x ^= x >> 32;
    x = (x * g.m3) & kMask63;
    x ^= x >> 30;
    return x;
}

__host__ __device__ __forceinline__ uint64_t shuffle_step(const GenParams &g, uint64_t x) {
    x = (x * g.sh_o1 + g.sh_c1) & g.sh_mask;
    x ^= x >> g.sh_r;
    x = (x * g.sh_o2) & g.sh_mask;
    x ^= x >> g.sh_r;
    return x;
}

// Inverse of shuffle_step on [0, 2^m): x ^= x >> r is its own inverse when
// 2r >= m (the shifted bits never reach the low half twice), and an odd
// multiplier is inverted by its inverse modulo 2^64.
__host__ __device__ __forceinline__ uint64_t shuffle_step_inv(const GenParams &g, uint64_t x) {
    x ^= x >> g.sh_r;
    x = (x * g.sh_i2) & g.sh_mask;
    x ^= x >> g.sh_r;
    return ((x - g.sh_c1) * g.sh_i1) & g.sh_mask;
}

// Inverse of an odd a modulo 2^64 by Newton's iteration (a * a == 1 mod 8,
// and each step doubles the number of correct low bits: 3, 6, 12, 24, 48, 96).
__host__ __device__ __forceinline__ uint64_t inverse_odd(uint64_t a) {
    uint64_t x = a;
    for (int i = 0; i < 5; ++i)
        x *= 2 - a * x;
    return x;
}

void gen_fill(hipStream_t s, const GenParams &g, int side, uint64_t rows, uint64_t *keys, const OutColSet &cols);
void gen_expected_checksum(hipStream_t s, const GenParams &g, unsigned long long *acc3);

} // namespace phjd
