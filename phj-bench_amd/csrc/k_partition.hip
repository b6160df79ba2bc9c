// k_partition.hip — radix partitioning on gfx950.
//
// Restates the behaviour of scatterBatch / radixShuffle / refineToLeaves
// (src/RadixPartition.cpp:316-393, :417-539, :546-577): rows are grouped by
// the hash bit window h >> (64 - consumed - bits) taken from the top of
// intHash64, one pass at a time, each pass splitting every segment of the
// previous pass. The CPU version appends to per-partition OutBlock chains; here
// every pass is histogram -> column scan -> scatter into one contiguous column
// per field, so partition p of pass k occupies a contiguous row range and no
// chains are needed.
//
// Kernels per pass:
//   k_seg_tiles     tiles and scan chunks per input segment (one workgroup)
//   k_hist          per-tile digit counts from per-wavefront LDS sub-histograms,
//                   stored tile-major (one contiguous row of 2^bits per tile)
//   k_colscan_chunk per (segment, chunk of tiles): exclusive column scan of the
//                   tile rows, chunk totals
//   k_colscan_seg   per segment: chunk bases, digit bases, output segment offsets
//   k_scatter       stable ranking inside the tile (wave ballots per digit bit,
//                   mbcnt, per-wave digit counts), then each column is staged
//                   in LDS in digit order and leaves as contiguous runs
//
// Every pass is stable: inside a tile rows keep their input order within a
// digit, tiles take their digit bases in tile order and segments in segment
// order, so partition p holds its rows in input order. The reference's
// radixShuffle yields the same order at any thread count: each thread
// scatters a contiguous run of blocks in order and the per-partition chains
// are concatenated in thread order (src/RadixPartition.cpp:144-205, :316-393).
#include <type_traits>

#include "kernels.h"

namespace phjd {

namespace {

constexpr int kPartThreads = 512;
constexpr int kPartRpt = 8;
constexpr int kTile = kPartThreads * kPartRpt; // rows per tile
// The scatter covers the same tile with twice the threads and half the rows
// each: at 64 VGPRs two of its workgroups share a CU (eight waves per SIMD),
// where 512 threads x 8 rows needed 146 VGPRs and ran one workgroup per CU.
constexpr int kScatThreads = 1024;
constexpr int kScatRpt = kTile / kScatThreads;
constexpr int kMaxDigits = 1 << kDeviceMaxPassBits; // digits of one device pass
constexpr int kChunkTiles = 256;               // tiles per column-scan chunk
constexpr int kScanThreads = 512;
constexpr int kScanPer = 8;
constexpr int kScanBlock = kScanThreads * kScanPer;

struct TileLoc {
    uint32_t tile, seg, local, ntiles, first;
    uint32_t row0, len;
};

// seg_first holds, for each segment, the first tile (entries 0..n_seg) and then
// the first scan chunk (entries n_seg+1 .. 2*n_seg+1).
__device__ __forceinline__ bool locate_tile(uint32_t t, const uint32_t *seg_off, const uint32_t *seg_first, uint32_t n_seg,
                                            TileLoc &L) {
    if (t >= seg_first[n_seg])
        return false;
    const uint32_t seg = upper_bound_idx(seg_first, n_seg, t);
    L.tile = t;
    L.seg = seg;
    L.first = seg_first[seg];
    L.ntiles = seg_first[seg + 1] - L.first;
    L.local = t - L.first;
    L.row0 = seg_off[seg] + L.local * kTile;
    const uint32_t end = seg_off[seg + 1];
    L.len = min(static_cast<uint32_t>(kTile), end - L.row0);
    return true;
}

__device__ __forceinline__ uint32_t n_tiles_of(uint32_t rows) { return (rows + kTile - 1) / kTile; }
__device__ __forceinline__ uint32_t n_chunks_of(uint32_t tiles) { return (tiles + kChunkTiles - 1) / kChunkTiles; }

// Bit window of segment s: the pass's uniform (shift, bits), or per segment
// from sb[s] = shift | bits << 8 (segments of batched joins split by their own
// bit counts). Output segment d of s is out_first[s] + d (uniform: s * B).
__device__ __forceinline__ void seg_window(const PassBits &p, uint32_t s, uint32_t &shift, uint32_t &bits) {
    if (p.sb) {
        const uint32_t v = p.sb[s];
        shift = v & 0xFFu;
        bits = v >> 8;
    } else {
        shift = p.shift;
        bits = p.bits;
    }
}

__device__ __forceinline__ uint32_t digit_of(uint64_t h, uint32_t shift, uint32_t bits) {
    return bits ? static_cast<uint32_t>(h >> shift) & ((1u << bits) - 1u) : 0u;
}

__device__ __forceinline__ uint32_t out_first_of(const PassBits &p, uint32_t s) {
    return p.out_first ? p.out_first[s] : s << p.bits;
}

// Tiles and chunks per segment, exclusively scanned (single workgroup).
__global__ __launch_bounds__(1024) void k_seg_tiles(const uint32_t *__restrict__ seg_off, uint32_t n_seg,
                                                    uint32_t *__restrict__ seg_first) {
    __shared__ uint32_t tmp[1024 / 64 + 1];
    const uint32_t per = (n_seg + 1023) / 1024;
    const uint32_t a = threadIdx.x * per;
    const uint32_t b = min(n_seg, a + per);
    uint32_t nt = 0, nc = 0;
    for (uint32_t s = a; s < b; ++s) {
        const uint32_t t = n_tiles_of(seg_off[s + 1] - seg_off[s]);
        nt += t;
        nc += n_chunks_of(t);
    }
    uint32_t tot_t, tot_c;
    uint32_t ot = block_excl_scan<1024>(nt, tmp, &tot_t);
    uint32_t oc = block_excl_scan<1024>(nc, tmp, &tot_c);
    uint32_t *chunk_first = seg_first + n_seg + 1;
    for (uint32_t s = a; s < b; ++s) {
        seg_first[s] = ot;
        chunk_first[s] = oc;
        const uint32_t t = n_tiles_of(seg_off[s + 1] - seg_off[s]);
        ot += t;
        oc += n_chunks_of(t);
    }
    if (threadIdx.x == 0) {
        seg_first[n_seg] = tot_t;
        chunk_first[n_seg] = tot_c;
    }
}

__global__ void k_seg_single(uint32_t *seg_off, uint32_t rows) {
    seg_off[0] = 0;
    seg_off[1] = rows;
}

// Per-tile digit counts, hist[tile * B + d]. With DIG the pass's digit of
// every row comes from `digits` (one byte per row, written by the previous
// pass's scatter) instead of hashing the keys: 1 byte read per row, not 8.
template <bool DIG>
__global__ __launch_bounds__(kPartThreads) void k_hist(const uint64_t *__restrict__ keys, const uint8_t *__restrict__ digits,
                                                       const uint32_t *__restrict__ seg_off,
                                                       const uint32_t *__restrict__ seg_first, uint32_t n_seg,
                                                       PassBits pb, uint32_t *__restrict__ hist) {
    constexpr int W = kPartThreads / 64;
    __shared__ uint32_t sub[W][kMaxDigits];
    TileLoc L;
    if (!locate_tile(xcd_share(blockIdx.x, seg_first[n_seg]), seg_off, seg_first, n_seg, L))
        return;
    const uint32_t B = 1u << pb.bits;
    uint32_t shift, bits;
    seg_window(pb, L.seg, shift, bits);
    for (uint32_t i = threadIdx.x; i < W * kMaxDigits; i += kPartThreads)
        (&sub[0][0])[i] = 0;
    __syncthreads();
    const uint32_t wave = threadIdx.x / 64;
    // all loads first: the LDS atomics below would otherwise order each load
    // after the previous row's atomic
    if constexpr (DIG) {
        uint32_t dg[kPartRpt];
#pragma unroll
        for (int r = 0; r < kPartRpt; ++r) {
            const uint32_t idx = r * kPartThreads + threadIdx.x;
            dg[r] = __builtin_nontemporal_load(digits + L.row0 + min(idx, L.len - 1)); // past the tile: masked below
        }
#pragma unroll
        for (int r = 0; r < kPartRpt; ++r) {
            const uint32_t idx = r * kPartThreads + threadIdx.x;
            if (idx < L.len)
                atomicAdd(&sub[wave][dg[r]], 1u);
        }
    } else {
        uint64_t k[kPartRpt];
#pragma unroll
        for (int r = 0; r < kPartRpt; ++r) {
            const uint32_t idx = r * kPartThreads + threadIdx.x;
            k[r] = __builtin_nontemporal_load(keys + L.row0 + min(idx, L.len - 1)); // past the tile: masked below
        }
#pragma unroll
        for (int r = 0; r < kPartRpt; ++r) {
            const uint32_t idx = r * kPartThreads + threadIdx.x;
            if (idx < L.len)
                atomicAdd(&sub[wave][digit_of(int_hash64(k[r]), shift, bits)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < B; d += kPartThreads) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < W; ++w)
            c += sub[w][d];
        hist[L.tile * B + d] = c;
    }
}

// Exclusive scan of column d of rows [r0, r1) of a row-major table with B
// columns, in place, starting from `run`; returns the column total. Loads are
// issued kColUnroll rows at a time, so a column walk has that many reads in
// flight instead of one.
constexpr int kColUnroll = 8;
__device__ __forceinline__ uint32_t excl_scan_column(uint32_t *__restrict__ t, uint32_t r0, uint32_t r1, uint32_t B,
                                                     uint32_t d, uint32_t run) {
    uint32_t r = r0;
    for (; r + kColUnroll <= r1; r += kColUnroll) {
        uint32_t v[kColUnroll];
#pragma unroll
        for (int i = 0; i < kColUnroll; ++i)
            v[i] = t[(r + i) * B + d];
#pragma unroll
        for (int i = 0; i < kColUnroll; ++i) {
            t[(r + i) * B + d] = run;
            run += v[i];
        }
    }
    for (; r < r1; ++r) {
        const uint32_t v = t[r * B + d];
        t[r * B + d] = run;
        run += v;
    }
    return run;
}

// Exclusive column scan of the tile rows of one chunk (within one segment);
// the chunk's per-digit totals go to chunk_tot[chunk * B + d].
__global__ __launch_bounds__(kMaxDigits) void k_colscan_chunk(const uint32_t *__restrict__ seg_first, uint32_t n_seg,
                                                              uint32_t bits, uint32_t *__restrict__ hist,
                                                              uint32_t *__restrict__ chunk_tot) {
    const uint32_t *chunk_first = seg_first + n_seg + 1;
    const uint32_t c = blockIdx.x;
    if (c >= chunk_first[n_seg])
        return;
    const uint32_t B = 1u << bits;
    const uint32_t s = upper_bound_idx(chunk_first, n_seg, c);
    const uint32_t k = c - chunk_first[s];
    const uint32_t t0 = seg_first[s] + k * kChunkTiles;
    const uint32_t t1 = min(seg_first[s + 1], t0 + kChunkTiles);
    const uint32_t d = threadIdx.x;
    if (d >= B)
        return;
    chunk_tot[c * B + d] = excl_scan_column(hist, t0, t1, B, d, 0u);
}

// Per segment: turn chunk totals into chunk bases (absolute row positions),
// and write the output segment offsets out_off[s * B + d].
__global__ __launch_bounds__(kMaxDigits) void k_colscan_seg(const uint32_t *__restrict__ seg_off,
                                                            const uint32_t *__restrict__ seg_first, uint32_t n_seg,
                                                            PassBits pb, uint32_t *__restrict__ chunk_tot,
                                                            uint32_t *__restrict__ out_off) {
    __shared__ uint32_t tmp[kMaxDigits / 64 + 1];
    const uint32_t *chunk_first = seg_first + n_seg + 1;
    const uint32_t s = blockIdx.x;
    const uint32_t B = 1u << pb.bits;
    uint32_t shift_s, bits_s;
    seg_window(pb, s, shift_s, bits_s);
    const uint32_t Bs = 1u << bits_s; // digits of this segment (the rest count zero)
    const uint32_t d = threadIdx.x;
    const uint32_t c0 = chunk_first[s], c1 = chunk_first[s + 1];
    const uint32_t run = d < B ? excl_scan_column(chunk_tot, c0, c1, B, d, 0u) : 0u;
    uint32_t total;
    const uint32_t base = seg_off[s] + block_excl_scan<kMaxDigits>(run, tmp, &total);
    if (d < B) {
        uint32_t c = c0;
        for (; c + kColUnroll <= c1; c += kColUnroll) {
            uint32_t v[kColUnroll];
#pragma unroll
            for (int i = 0; i < kColUnroll; ++i)
                v[i] = chunk_tot[(c + i) * B + d];
#pragma unroll
            for (int i = 0; i < kColUnroll; ++i)
                chunk_tot[(c + i) * B + d] = v[i] + base;
        }
        for (; c < c1; ++c)
            chunk_tot[c * B + d] += base;
    }
    if (d < Bs)
        out_off[out_first_of(pb, s) + d] = base;
    if (s == n_seg - 1 && d == 0)
        out_off[out_first_of(pb, n_seg)] = seg_off[n_seg];
}

// ---- generic device-wide exclusive scan of uint32 (three kernels) ----
__device__ __forceinline__ uint64_t scan_count(uint64_t max_n, const uint32_t *count_ptr, uint32_t mul) {
    return count_ptr ? static_cast<uint64_t>(*count_ptr) * mul : max_n;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *__restrict__ data, uint64_t max_n,
                                                              const uint32_t *count_ptr, uint32_t mul,
                                                              uint32_t *__restrict__ block_sums) {
    __shared__ uint32_t tmp[kScanThreads / 64 + 1];
    const uint64_t n = scan_count(max_n, count_ptr, mul);
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanBlock;
    if (base >= n)
        return;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const uint64_t i = base + static_cast<uint64_t>(k) * kScanThreads + threadIdx.x;
        if (i < n)
            s += data[i];
    }
    uint32_t total;
    block_excl_scan<kScanThreads>(s, tmp, &total);
    if (threadIdx.x == 0)
        block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t *__restrict__ block_sums, uint64_t max_n,
                                                   const uint32_t *count_ptr, uint32_t mul) {
    __shared__ uint32_t tmp[1024 / 64 + 1];
    const uint64_t n = scan_count(max_n, count_ptr, mul);
    const uint32_t nb = static_cast<uint32_t>((n + kScanBlock - 1) / kScanBlock);
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? block_sums[i] : 0u;
        uint32_t total;
        const uint32_t off = block_excl_scan<1024>(v, tmp, &total);
        if (i < nb)
            block_sums[i] = carry + off;
        carry += total;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(uint32_t *__restrict__ data, uint64_t max_n,
                                                            const uint32_t *count_ptr, uint32_t mul,
                                                            const uint32_t *__restrict__ block_sums) {
    __shared__ uint32_t tmp[kScanThreads / 64 + 1];
    const uint64_t n = scan_count(max_n, count_ptr, mul);
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanBlock;
    if (base >= n)
        return;
    const uint64_t my = base + static_cast<uint64_t>(threadIdx.x) * kScanPer;
    uint32_t v[kScanPer];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = (my + k < n) ? data[my + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t off = block_excl_scan<kScanThreads>(s, tmp, &total) + block_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (my + k < n)
            data[my + k] = off;
        off += v[k];
    }
}

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

template <bool NT, class T> __device__ __forceinline__ T ld(const T *p) {
    if constexpr (!NT) {
        return *p;
    } else if constexpr (sizeof(T) == 16) {
        const u64x2 v = __builtin_nontemporal_load(reinterpret_cast<const u64x2 *>(p));
        T r;
        __builtin_memcpy(&r, &v, 16);
        return r;
    } else {
        return __builtin_nontemporal_load(p);
    }
}

template <bool NT, class T> __device__ __forceinline__ void st(T *p, const T &v) {
    if constexpr (!NT) {
        *p = v;
    } else if constexpr (sizeof(T) == 16) {
        u64x2 x;
        __builtin_memcpy(&x, &v, 16);
        __builtin_nontemporal_store(x, reinterpret_cast<u64x2 *>(p));
    } else {
        __builtin_nontemporal_store(v, p);
    }
}

constexpr int kScatWaves = kScatThreads / 64;
constexpr int kWaveRows = kTile / kScatWaves; // rows of a tile ranked by one wave

struct ScatterShared {
    union {
        uint64_t stage[kTile];
        // per-wave digit counts while ranking, then each wave's base per digit
        uint16_t wcount[kScatWaves][kMaxDigits];
    };
    uint16_t slot_digit[kTile];
    uint32_t cursor[kMaxDigits];
    uint32_t lstart[kMaxDigits];
    uint32_t gbase[kMaxDigits];
    uint32_t scan_tmp[kScatThreads / 64 + 1];
};
static_assert(sizeof(uint16_t) * kScatWaves * kMaxDigits <= sizeof(uint64_t) * kTile, "wave counts fit the stage");
static_assert(kWaveRows < 65536, "a wave's digit count fits 16 bits");

// Row of the tile held by thread tid in its r-th register. ATOM (atomic
// ranking): rows r * kScatThreads + tid. Otherwise wave w holds rows
// [w * kWaveRows, (w + 1) * kWaveRows), 64 consecutive rows per register.
// Either way one wave's load touches 64 consecutive rows.
template <bool ATOM> __device__ __forceinline__ uint32_t tile_row(int r, uint32_t tid) {
    if constexpr (ATOM)
        return r * kScatThreads + tid;
    else
        return (tid >> 6) * kWaveRows + r * 64 + (tid & 63u);
}

constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

// Values already in registers are staged in LDS at their digit-ordered slot,
// then leave in slot order so each digit's rows form one contiguous run.
template <bool NTS, class T>
__device__ __forceinline__ void write_staged(ScatterShared &sh, const T (&v)[kScatRpt], T *__restrict__ out,
                                             const TileLoc &L, const uint32_t (&slot)[kScatRpt]) {
    constexpr uint32_t kPerRound = sizeof(T) <= 8 ? kTile : kTile * 8 / sizeof(T);
    constexpr uint32_t kRounds = kTile / kPerRound;
    T *stg = reinterpret_cast<T *>(sh.stage);
#pragma unroll
    for (uint32_t round = 0; round < kRounds; ++round) {
        const uint32_t lo = round * kPerRound;
#pragma unroll
        for (int r = 0; r < kScatRpt; ++r)
            if (slot[r] - lo < kPerRound) // kNoSlot never qualifies
                stg[slot[r] - lo] = v[r];
        __syncthreads();
        const uint32_t hi = min(L.len, lo + kPerRound);
        for (uint32_t k = lo + threadIdx.x; k < hi; k += kScatThreads) {
            const uint32_t d = sh.slot_digit[k];
            st<NTS>(out + sh.gbase[d] + (k - sh.lstart[d]), stg[k - lo]);
        }
        __syncthreads();
    }
}

template <bool NTL, bool NTS, bool ATOM, class T>
__device__ __forceinline__ void scatter_column(ScatterShared &sh, const T *__restrict__ in, T *__restrict__ out,
                                               const TileLoc &L, const uint32_t (&slot)[kScatRpt]) {
    // the row indices are recomputed per column from an opaque copy of the
    // thread id: hoisted out of the column loop they would stay live in
    // registers across every column and spill
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    T v[kScatRpt];
#pragma unroll
    for (int r = 0; r < kScatRpt; ++r) {
        const uint32_t idx = tile_row<ATOM>(r, tid);
        v[r] = ld<NTL>(in + L.row0 + min(idx, L.len - 1)); // rows past the tile carry no slot
    }
    write_staged<NTS, T>(sh, v, out, L, slot);
}

// Stable rank of each of a wave's rows among the tile's rows of its digit
// that precede it in the wave's range: lanes holding the same digit find each
// other with one ballot per digit bit, the rank inside a 64-row step is the
// count of such lanes below (mbcnt), and the wave's running count per digit
// (its own LDS row, so no atomics) carries across its steps. The lowest lane
// of each digit advances the count. Returns rank << kDeviceMaxPassBits | digit.
__device__ __forceinline__ uint32_t wave_rank(uint16_t *__restrict__ wcnt, uint32_t d, bool valid, uint32_t bits) {
    uint64_t same = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < kDeviceMaxPassBits; ++b) {
        if (b < bits) { // uniform over the workgroup
            const bool one = (d >> b) & 1u;
            const uint64_t m = __ballot(one);
            same &= one ? m : ~m;
        }
    }
    const uint32_t below = mbcnt(same);
    const uint32_t base = wcnt[d];
    if (valid && below == 0)
        wcnt[d] = static_cast<uint16_t>(base + __popcll(same));
    return ((base + below) << kDeviceMaxPassBits) | d;
}

// With DIG the scatter also writes the next pass's digit of every row
// (hash bits [next_shift, next_shift + next_bits), at most 8) as a byte
// column in output order, so the next pass's histogram need not re-read and
// re-hash the keys.
//
// ATOM ranks rows with one LDS atomic each instead (not stable; kept as a
// tuning variant for comparison).
template <bool NTL, bool NTS, bool DIG, bool ATOM>
__global__ __launch_bounds__(kScatThreads) __attribute__((amdgpu_waves_per_eu(8)))
void k_scatter(const uint64_t *__restrict__ keys_in, ColSet cin,
                                                          uint64_t *__restrict__ keys_out, OutColSet cout,
                                                          const uint32_t *__restrict__ seg_off,
                                                          const uint32_t *__restrict__ seg_first, uint32_t n_seg,
                                                          PassBits pb,
                                                          const uint32_t *__restrict__ hist,
                                                          const uint32_t *__restrict__ chunk_base,
                                                          uint8_t *__restrict__ digits_out, uint32_t next_shift,
                                                          uint32_t next_bits) {
    __shared__ ScatterShared sh;
    TileLoc L;
    if (!locate_tile(xcd_share(blockIdx.x, seg_first[n_seg]), seg_off, seg_first, n_seg, L))
        return;
    const uint32_t B = 1u << pb.bits;
    uint32_t shift, bits;
    seg_window(pb, L.seg, shift, bits);
    if constexpr (ATOM) {
        for (uint32_t i = threadIdx.x; i < B; i += kScatThreads)
            sh.cursor[i] = 0;
    } else {
        uint32_t *w = reinterpret_cast<uint32_t *>(&sh.wcount[0][0]);
        for (uint32_t i = threadIdx.x; i < kScatWaves * kMaxDigits / 2; i += kScatThreads)
            w[i] = 0;
    }
    __syncthreads();

    // Rank each row inside its digit (rank << kDeviceMaxPassBits | digit).
    uint64_t key[kScatRpt];
    uint32_t slot[kScatRpt];
#pragma unroll
    for (int r = 0; r < kScatRpt; ++r) {
        const uint32_t idx = tile_row<ATOM>(r, threadIdx.x);
        key[r] = ld<NTL>(keys_in + L.row0 + min(idx, L.len - 1)); // past the tile: masked below
    }
    const uint32_t wave = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < kScatRpt; ++r) {
        const uint32_t idx = tile_row<ATOM>(r, threadIdx.x);
        const uint32_t d = digit_of(int_hash64(key[r]), shift, bits);
        if constexpr (ATOM) {
            slot[r] = idx < L.len ? (atomicAdd(&sh.cursor[d], 1u) << kDeviceMaxPassBits) | d : kNoSlot;
        } else {
            const uint32_t s = wave_rank(sh.wcount[wave], d, idx < L.len, bits);
            slot[r] = idx < L.len ? s : kNoSlot;
        }
    }
    __syncthreads();
    if constexpr (!ATOM) {
        // wave counts -> each wave's base inside its digit; digit totals
        if (threadIdx.x < B) {
            uint32_t acc = 0;
#pragma unroll
            for (int w = 0; w < kScatWaves; ++w) {
                const uint32_t c = sh.wcount[w][threadIdx.x];
                sh.wcount[w][threadIdx.x] = static_cast<uint16_t>(acc);
                acc += c;
            }
            sh.cursor[threadIdx.x] = acc;
        }
        __syncthreads();
    }
    {
        const uint32_t c = threadIdx.x < B ? sh.cursor[threadIdx.x] : 0u;
        uint32_t total;
        const uint32_t off = block_excl_scan<kScatThreads>(c, sh.scan_tmp, &total);
        if (threadIdx.x < B) {
            const uint32_t chunk = seg_first[n_seg + 1 + L.seg] + L.local / kChunkTiles;
            sh.lstart[threadIdx.x] = off;
            sh.gbase[threadIdx.x] = hist[L.tile * B + threadIdx.x] + chunk_base[chunk * B + threadIdx.x];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kScatRpt; ++r) {
        if (slot[r] != kNoSlot) {
            const uint32_t d = slot[r] & (kMaxDigits - 1u);
            uint32_t sl = sh.lstart[d] + (slot[r] >> kDeviceMaxPassBits);
            if constexpr (!ATOM)
                sl += sh.wcount[wave][d];
            sh.slot_digit[sl] = static_cast<uint16_t>(d);
            slot[r] = sl;
        }
    }
    __syncthreads(); // the wave bases share LDS with the staging area below

    write_staged<NTS, uint64_t>(sh, key, keys_out, L, slot);
    if constexpr (DIG) {
        // the next pass's digits, hashed again here rather than held in
        // registers through the ranking
        uint8_t next[kScatRpt];
#pragma unroll
        for (int r = 0; r < kScatRpt; ++r)
            next[r] = static_cast<uint8_t>(digit_of(int_hash64(key[r]), next_shift, next_bits));
        write_staged<NTS, uint8_t>(sh, next, digits_out, L, slot);
    }
    for (uint32_t c = 0; c < cin.n; ++c) {
        switch (col_type(cin, c)) {
        case 0:
            scatter_column<NTL, NTS, ATOM, uint8_t>(sh, static_cast<const uint8_t *>(col_ptr(cin, c)), static_cast<uint8_t *>(col_ptr(cout, c)),
                                    L, slot);
            break;
        case 1:
            scatter_column<NTL, NTS, ATOM, uint16_t>(sh, static_cast<const uint16_t *>(col_ptr(cin, c)),
                                     static_cast<uint16_t *>(col_ptr(cout, c)), L, slot);
            break;
        case 2:
            scatter_column<NTL, NTS, ATOM, uint32_t>(sh, static_cast<const uint32_t *>(col_ptr(cin, c)),
                                     static_cast<uint32_t *>(col_ptr(cout, c)), L, slot);
            break;
        case 3:
            scatter_column<NTL, NTS, ATOM, uint64_t>(sh, static_cast<const uint64_t *>(col_ptr(cin, c)),
                                     static_cast<uint64_t *>(col_ptr(cout, c)), L, slot);
            break;
        case kTypeRowIndex: {
            // row indices of the input, synthesised (nothing to read)
            uint32_t v[kScatRpt];
#pragma unroll
            for (int r = 0; r < kScatRpt; ++r)
                v[r] = L.row0 + min(tile_row<ATOM>(r, threadIdx.x), L.len - 1);
            write_staged<NTS, uint32_t>(sh, v, static_cast<uint32_t *>(col_ptr(cout, c)), L, slot);
            break;
        }
        default:
            scatter_column<NTL, NTS, ATOM, ulonglong2>(sh, static_cast<const ulonglong2 *>(col_ptr(cin, c)),
                                       static_cast<ulonglong2 *>(col_ptr(cout, c)), L, slot);
            break;
        }
    }
}

__global__ void k_u32_to_u64(const uint32_t *__restrict__ in, uint64_t *__restrict__ out, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = in[i];
}

} // namespace

uint32_t radix_tile_rows() { return kTile; }

uint64_t radix_max_tiles(uint64_t rows, uint32_t n_seg) { return (rows + kTile - 1) / kTile + n_seg; }

uint64_t radix_max_chunks(uint64_t rows, uint32_t n_seg) { return radix_max_tiles(rows, n_seg) / kChunkTiles + n_seg; }

uint64_t radix_scan_blocks(uint64_t elems) { return (elems + kScanBlock - 1) / kScanBlock; }

void scan_u32(hipStream_t s, uint32_t *data, uint64_t max_n, const uint32_t *count_ptr, uint32_t mul,
              uint32_t *block_sums, KernelTimer *tm, const char *name) {
    const uint64_t nb = radix_scan_blocks(max_n);
    if (nb == 0)
        return;
    TimedScope ts(tm, name, s);
    k_scan_reduce<<<static_cast<uint32_t>(nb), kScanThreads, 0, s>>>(data, max_n, count_ptr, mul, block_sums);
    k_scan_top<<<1, 1024, 0, s>>>(block_sums, max_n, count_ptr, mul);
    k_scan_down<<<static_cast<uint32_t>(nb), kScanThreads, 0, s>>>(data, max_n, count_ptr, mul, block_sums);
}

void launch_seg_single(hipStream_t s, uint32_t *seg_off, uint32_t rows) { k_seg_single<<<1, 1, 0, s>>>(seg_off, rows); }

void launch_u32_to_u64(hipStream_t s, const uint32_t *in, uint64_t *out, uint64_t n) {
    if (n == 0)
        return;
    k_u32_to_u64<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, s>>>(in, out, n);
}

void radix_pass(hipStream_t s, const uint64_t *keys_in, const ColSet &cin, uint64_t *keys_out, const OutColSet &cout,
                const uint32_t *seg_off, uint32_t n_seg, uint64_t rows, const PassBits &pb,
                uint32_t *out_seg_off, const RadixPassScratch &scr, uint32_t flags, KernelTimer *tm,
                const PassDigits &dig) {
    const uint32_t bits = pb.bits;
    const uint64_t max_tiles = radix_max_tiles(rows, n_seg);
    const uint64_t max_chunks = radix_max_chunks(rows, n_seg);
    {
        TimedScope ts(tm, "radix_plan", s);
        k_seg_tiles<<<1, 1024, 0, s>>>(seg_off, n_seg, scr.seg_tile_first);
    }
    if (rows > 0) {
        TimedScope ts(tm, "radix_hist", s);
        if (dig.in)
            k_hist<true><<<grid8(max_tiles), kPartThreads, 0, s>>>(keys_in, dig.in, seg_off, scr.seg_tile_first, n_seg, pb,
                                                                    scr.hist);
        else
            k_hist<false><<<grid8(max_tiles), kPartThreads, 0, s>>>(keys_in, nullptr, seg_off, scr.seg_tile_first, n_seg,
                                                                     pb, scr.hist);
    }
    {
        TimedScope ts(tm, "radix_scan", s);
        if (rows > 0)
            k_colscan_chunk<<<static_cast<uint32_t>(max_chunks), kMaxDigits, 0, s>>>(scr.seg_tile_first, n_seg, bits, scr.hist,
                                                                                      scr.block_sums);
        k_colscan_seg<<<n_seg, kMaxDigits, 0, s>>>(seg_off, scr.seg_tile_first, n_seg, pb, scr.block_sums, out_seg_off);
    }
    if (rows > 0) {
        TimedScope ts(tm, "radix_scatter", s);
        const bool ntl = flags & kRadixNtLoads, nts = flags & kRadixNtStores;
        auto pick = [&](auto atom) {
            constexpr bool A = decltype(atom)::value;
            return dig.out ? (ntl ? (nts ? k_scatter<true, true, true, A> : k_scatter<true, false, true, A>)
                                  : (nts ? k_scatter<false, true, true, A> : k_scatter<false, false, true, A>))
                           : (ntl ? (nts ? k_scatter<true, true, false, A> : k_scatter<true, false, false, A>)
                                  : (nts ? k_scatter<false, true, false, A> : k_scatter<false, false, false, A>));
        };
        auto k = (flags & kRadixAtomicRank) ? pick(std::true_type{}) : pick(std::false_type{});
        k<<<grid8(max_tiles), kScatThreads, 0, s>>>(keys_in, cin, keys_out, cout, seg_off, scr.seg_tile_first, n_seg, pb,
                                                     scr.hist, scr.block_sums, dig.out, dig.next_shift, dig.next_bits);
    }
}

} // namespace phjd
