// k_gen.hip — device-side synthetic workloads and their closed-form checksum.
#include "gen.h"

namespace phjd {

namespace {

__device__ __forceinline__ uint64_t gen_build_key(const GenParams &g, uint64_t i) {
    if (g.dist == 1)
        return int_hash64((g.seed ^ kRefBuildKeySalt) + i); // src/DataGen.cpp:86,94
    if (g.dist == 2) {
        uint64_t y = 0;
        if (g.sh_mask != 0) {
            y = shuffle_step(g, i);
            while (y >= g.build_total)
                y = shuffle_step(g, y);
        }
        const uint64_t d = upper_bound_idx(g.prefix, static_cast<uint32_t>(g.build_rows + 1), y);
        return perm63(g, d + 1);
    }
    return perm63(g, i + 1);
}

// Probe key at row j; pick = matched build row (or distinct key for Zipf),
// ~0 for a miss.
__device__ __forceinline__ uint64_t gen_probe_key(const GenParams &g, uint64_t j, uint64_t &pick) {
    if (g.dist == 1) {
        pick = int_hash64((g.seed ^ kRefProbePickSalt) + j) % g.build_rows; // src/DataGen.cpp:124-126
        return int_hash64((g.seed ^ kRefBuildKeySalt) + pick);
    }
    if (mulhi64(rand64(g.b_sel, j), 1000000ull) >= g.match_ppm) {
        pick = ~0ull;
        return (1ull << 63) | (rand64(g.b_miss, j) >> 1);
    }
    if (g.dist == 2) {
        const uint64_t u = rand64(g.b_rank, j) >> 11;
        // number of thresholds <= u
        uint32_t lo = 0, hi = g.n_rank;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (g.rank_t[mid] <= u)
                lo = mid + 1;
            else
                hi = mid;
        }
        pick = lo;
        return perm63(g, pick + 1);
    }
    pick = mulhi64(rand64(g.b_pick, j), g.build_rows);
    return perm63(g, pick + 1);
}

__device__ __forceinline__ void gen_payload(const GenParams &g, int side, uint32_t c, uint64_t i, uint64_t &lo, uint64_t &hi) {
    const uint32_t type = side == 0 ? g.build_types[c] : g.probe_types[c];
    if (g.dist == 1) {
        // src/DataGen.cpp:95-101, :129-135
        const uint64_t pay_salt = g.seed ^ (side == 0 ? kRefBuildPaySalt : kRefProbePaySalt);
        const uint64_t col_salt = pay_salt ^ (kRefColumnSaltStride * (c + 1));
        lo = int_hash64(col_salt ^ i);
        hi = type == 4 ? int_hash64(col_salt ^ (i + (1ull << 33))) : 0;
        return;
    }
    lo = rand64(side == 0 ? g.b_bpay[c] : g.b_ppay[c], i);
    hi = type == 4 ? rand64(side == 0 ? g.b_bpayhi[c] : g.b_ppayhi[c], i) : 0;
}

__global__ __launch_bounds__(256) void k_gen(GenParams g, int side, uint64_t rows, uint64_t *__restrict__ keys, OutColSet cols) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < rows; i += stride) {
        uint64_t pick;
        keys[i] = side == 0 ? gen_build_key(g, i) : gen_probe_key(g, i, pick);
        for (uint32_t c = 0; c < cols.n; ++c) {
            uint64_t lo, hi;
            gen_payload(g, side, c, i, lo, hi);
            store_elem(cols.ptr[c], cols.types[c], i, lo, hi);
        }
    }
}

// Build row whose shuffled slot is y (Zipf): the inverse cycle walk of
// gen_build_key's forward walk (the forward walk steps until it lands below
// build_total, so the inverse steps back until it does).
__device__ __forceinline__ uint64_t zipf_row_of_slot(const GenParams &g, uint64_t y) {
    if (g.sh_mask == 0)
        return 0;
    uint64_t x = shuffle_step_inv(g, y);
    while (x >= g.build_total)
        x = shuffle_step_inv(g, x);
    return x;
}

__device__ __forceinline__ void digest_row(const GenParams &g, uint32_t row_size, uint64_t key, uint64_t build_row,
                                           uint64_t j, uint64_t &s0, uint64_t &s1) {
    Packer p(row_size);
    p.push_le8(key, 8);
    for (uint32_t c = 0; c < g.n_build_cols; ++c) {
        uint64_t lo, hi;
        gen_payload(g, 0, c, build_row, lo, hi);
        p.push(lo, hi, type_size(g.build_types[c]));
    }
    for (uint32_t c = 0; c < g.n_probe_cols; ++c) {
        uint64_t lo, hi;
        gen_payload(g, 1, c, j, lo, hi);
        p.push(lo, hi, type_size(g.probe_types[c]));
    }
    p.finish();
    s0 += p.d.a;
    s1 += p.d.c;
}

// Order-independent checksum of the exact join output the generator implies.
// Uniform and reference streams: each matching probe row pairs with the one
// build row it was drawn from. Zipf: a probe row that drew distinct key d
// pairs with the k_d build rows of d, the rows whose shuffled slots are
// [prefix[d], prefix[d+1]).
__global__ __launch_bounds__(256) void k_gen_checksum(GenParams g, uint32_t row_size, unsigned long long *__restrict__ acc) {
    __shared__ unsigned long long part[3][4];
    uint64_t s0 = 0, s1 = 0, cnt = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < g.probe_rows; j += stride) {
        uint64_t pick;
        const uint64_t key = gen_probe_key(g, j, pick);
        if (pick == ~0ull || g.build_rows == 0)
            continue;
        if (g.dist == 2) {
            const uint64_t y0 = g.prefix[pick], y1 = g.prefix[pick + 1];
            for (uint64_t y = y0; y < y1; ++y)
                digest_row(g, row_size, key, zipf_row_of_slot(g, y), j, s0, s1);
            cnt += y1 - y0;
            continue;
        }
        digest_row(g, row_size, key, pick, j, s0, s1);
        cnt += 1;
    }
    s0 = wave_sum_u64(s0);
    s1 = wave_sum_u64(s1);
    cnt = wave_sum_u64(cnt);
    const uint32_t w = threadIdx.x / 64;
    if (lane_id() == 0) {
        part[0][w] = cnt;
        part[1][w] = s0;
        part[2][w] = s1;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        unsigned long long t = 0;
        for (int k = 0; k < 4; ++k)
            t += part[threadIdx.x][k];
        atomicAdd(&acc[threadIdx.x], t);
    }
}

} // namespace

void gen_fill(hipStream_t s, const GenParams &g, int side, uint64_t rows, uint64_t *keys, const OutColSet &cols) {
    if (rows == 0)
        return;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((rows + 255) / 256, 256 * 64));
    k_gen<<<grid, 256, 0, s>>>(g, side, rows, keys, cols);
}

void gen_expected_checksum(hipStream_t s, const GenParams &g, unsigned long long *acc3) {
    launch_fill_u64(s, reinterpret_cast<uint64_t *>(acc3), 3, 0);
    if (g.probe_rows == 0)
        return;
    uint32_t row_size = 8;
    for (uint32_t c = 0; c < g.n_build_cols; ++c)
        row_size += type_size(g.build_types[c]);
    for (uint32_t c = 0; c < g.n_probe_cols; ++c)
        row_size += type_size(g.probe_types[c]);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((g.probe_rows + 255) / 256, 256 * 16));
    k_gen_checksum<<<grid, 256, 0, s>>>(g, row_size, acc3);
}

} // namespace phjd
