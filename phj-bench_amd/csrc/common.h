// common.h — device helpers shared by the gfx950 kernels of libphj_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phjd {

constexpr uint32_t kInvalidRow = 0xFFFFFFFFu;        // empty RowRefCell marker (src/Types.h:14)
constexpr uint64_t kEmptyKey = 0xFFFFFFFFFFFFFFFFull; // unclaimed hash-table key slot
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kMask63 = 0x7FFFFFFFFFFFFFFFull;
constexpr int kMaxCols = 8;

// src/Hash.h:18-26 — intHash64 (xor salt, then Murmur3 fmix64).
__host__ __device__ __forceinline__ uint64_t int_hash64(uint64_t x) {
    x ^= 0x4cf5ad432745937full;
    x = (x ^ (x >> 33)) * 0xff51afd7ed558ccdull;
    x = (x ^ (x >> 33)) * 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// SplitMix64 output function; the harness generators and the row digest use it.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t rand64(uint64_t base, uint64_t i) { return mix64(base + (i + 1) * kGolden); }

__host__ __device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return static_cast<uint64_t>((static_cast<unsigned __int128>(a) * b) >> 64);
#endif
}

__host__ __device__ __forceinline__ uint32_t type_size(uint32_t t) { return t >= 4 ? 16u : (1u << t); }

// ---- wavefront helpers (64 lanes) ----
__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Number of set bits of `mask` in lanes below the calling lane (mbcnt).
__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Inclusive wave scan of a 32-bit value (64 lanes).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= static_cast<uint32_t>(d))
            v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_xor(v, d, 64);
    return v;
}

// Exclusive block scan of one value per thread; returns the thread's offset
// and writes the block total to *total. `lds` needs THREADS/64 + 1 words.
template <int THREADS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *lds, uint32_t *total) {
    constexpr int W = THREADS / 64;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63)
        lds[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < W; ++w) {
            const uint32_t t = lds[w];
            lds[w] = acc;
            acc += t;
        }
        lds[W] = acc;
    }
    __syncthreads();
    const uint32_t off = lds[wave] + inc - v;
    *total = lds[W];
    __syncthreads();
    return off;
}

// block_excl_scan of two values at once (the same barriers): `lds` needs
// 2 * (THREADS/64 + 1) words.
template <int THREADS>
__device__ __forceinline__ void block_excl_scan2(uint32_t v0, uint32_t v1, uint32_t *lds, uint32_t &o0, uint32_t &o1,
                                                 uint32_t &t0, uint32_t &t1) {
    constexpr int W = THREADS / 64;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    const uint32_t i0 = wave_incl_scan(v0);
    const uint32_t i1 = wave_incl_scan(v1);
    if (lane == 63) {
        lds[wave] = i0;
        lds[W + 1 + wave] = i1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a0 = 0, a1 = 0;
        for (int w = 0; w < W; ++w) {
            const uint32_t x0 = lds[w], x1 = lds[W + 1 + w];
            lds[w] = a0;
            lds[W + 1 + w] = a1;
            a0 += x0;
            a1 += x1;
        }
        lds[W] = a0;
        lds[2 * W + 1] = a1;
    }
    __syncthreads();
    o0 = lds[wave] + i0 - v0;
    o1 = lds[W + 1 + wave] + i1 - v1;
    t0 = lds[W];
    t1 = lds[2 * W + 1];
    __syncthreads();
}

// Largest i in [0, n) with a[i] <= x, for a non-decreasing array with a[0] <= x.
template <class T>
__device__ __forceinline__ uint32_t upper_bound_idx(const T *a, uint32_t n, T x) {
    uint32_t lo = 0, hi = n; // invariant: a[lo] <= x, answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Grid size rounded up to a multiple of the 8 XCDs of an MI355X.
__host__ __device__ __forceinline__ uint32_t grid8(uint64_t n) {
    return static_cast<uint32_t>(((n + 7) / 8) * 8);
}

// Work index of block b when only `total` of the launched blocks carry work
// (total known on the device, grid >= 8 * ceil(total / 8)). Workgroups are
// dispatched round-robin over the 8 XCDs, so block b runs on XCD b % 8; each
// XCD gets an equal, contiguous share of the work indices (neighbouring work
// items share that XCD's L2). Returns an index >= total for blocks without
// work.
__device__ __forceinline__ uint32_t xcd_share(uint32_t b, uint32_t total) {
    const uint32_t per = (total + 7) / 8;
    const uint32_t q = b / 8;
    return q < per ? (b % 8) * per + q : 0xFFFFFFFFu;
}

// ---- row digest (same definition as the oracle's) ----
struct Digest {
    uint64_t a, c;
    __host__ __device__ explicit Digest(uint64_t len) : a(0x243F6A8885A308D3ull ^ len), c(0x13198A2E03707344ull + len) {}
    __host__ __device__ __forceinline__ void word(uint64_t w) {
        a = mix64(a ^ w);
        c = mix64(c + w * kGolden);
    }
};

// Packs little-endian fields of 1..16 bytes into 8-byte words.
struct Packer {
    Digest d;
    uint64_t acc = 0;
    uint32_t fill = 0;
    __host__ __device__ explicit Packer(uint64_t len) : d(len) {}
    __host__ __device__ __forceinline__ void push_le8(uint64_t v, uint32_t sz) { // sz in 1..8, v zero above sz bytes
        acc |= fill == 0 ? v : (v << (8 * fill));
        const uint32_t nf = fill + sz;
        if (nf >= 8) {
            d.word(acc);
            acc = (fill == 0 || nf == 8) ? 0 : (v >> (8 * (8 - fill)));
            fill = nf - 8;
        } else {
            fill = nf;
        }
    }
    __host__ __device__ __forceinline__ void push(uint64_t lo, uint64_t hi, uint32_t sz) {
        if (sz <= 8) {
            push_le8(sz == 8 ? lo : (lo & ((1ull << (8 * sz)) - 1)), sz);
        } else {
            push_le8(lo, 8);
            push_le8(hi, 8);
        }
    }
    __host__ __device__ __forceinline__ void finish() {
        if (fill != 0)
            d.word(acc);
    }
};

// Column element load as (lo, hi) by type code.
__device__ __forceinline__ void load_elem(const void *col, uint32_t type, uint64_t i, uint64_t &lo, uint64_t &hi) {
    hi = 0;
    switch (type) {
    case 0: lo = static_cast<const uint8_t *>(col)[i]; break;
    case 1: lo = static_cast<const uint16_t *>(col)[i]; break;
    case 2: lo = static_cast<const uint32_t *>(col)[i]; break;
    case 3: lo = static_cast<const uint64_t *>(col)[i]; break;
    default: {
        const ulonglong2 v = static_cast<const ulonglong2 *>(col)[i];
        lo = v.x;
        hi = v.y;
    }
    }
}

__device__ __forceinline__ void store_elem(void *col, uint32_t type, uint64_t i, uint64_t lo, uint64_t hi) {
    switch (type) {
    case 0: static_cast<uint8_t *>(col)[i] = static_cast<uint8_t>(lo); break;
    case 1: static_cast<uint16_t *>(col)[i] = static_cast<uint16_t>(lo); break;
    case 2: static_cast<uint32_t *>(col)[i] = static_cast<uint32_t>(lo); break;
    case 3: static_cast<uint64_t *>(col)[i] = lo; break;
    default: static_cast<ulonglong2 *>(col)[i] = make_ulonglong2(lo, hi);
    }
}

// Column set passed by value to kernels.
struct ColSet {
    uint32_t n;
    uint32_t types[kMaxCols];
    const void *ptr[kMaxCols];
};

struct OutColSet {
    uint32_t n;
    uint32_t types[kMaxCols];
    void *ptr[kMaxCols];
};

// Field c of a by-value kernel argument with a wave-uniform c. Constant
// indices in every case keep the argument in the kernarg segment instead of
// copying it to scratch for a dynamic index.
template <class S>
__device__ __forceinline__ auto col_ptr(const S &s, uint32_t c) -> decltype(s.ptr[0]) {
    switch (c) {
    case 0: return s.ptr[0];
    case 1: return s.ptr[1];
    case 2: return s.ptr[2];
    case 3: return s.ptr[3];
    case 4: return s.ptr[4];
    case 5: return s.ptr[5];
    case 6: return s.ptr[6];
    default: return s.ptr[7];
    }
}

template <class S> __device__ __forceinline__ uint32_t col_type(const S &s, uint32_t c) {
    switch (c) {
    case 0: return s.types[0];
    case 1: return s.types[1];
    case 2: return s.types[2];
    case 3: return s.types[3];
    case 4: return s.types[4];
    case 5: return s.types[5];
    case 6: return s.types[6];
    default: return s.types[7];
    }
}

} // namespace phjd
