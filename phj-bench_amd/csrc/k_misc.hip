// k_misc.hip — streaming kernels: intHash64Batch, copy/fill, output checksum.
#include "kernels.h"

namespace phjd {

namespace {

// src/Hash.h:33-58 — out[i] = intHash64(in[i]); 16-byte vector loads/stores,
// grid-stride, purely HBM-bound (16 B per key).
__global__ __launch_bounds__(256) void k_hash_batch(const uint64_t *__restrict__ in, uint64_t n, uint64_t *__restrict__ out) {
    const uint64_t pairs = n / 2;
    const auto *in2 = reinterpret_cast<const ulonglong2 *>(in);
    auto *out2 = reinterpret_cast<ulonglong2 *>(out);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < pairs; i += stride) {
        const ulonglong2 v = in2[i];
        out2[i] = make_ulonglong2(int_hash64(v.x), int_hash64(v.y));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1))
        out[n - 1] = int_hash64(in[n - 1]);
}

__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ a, uint4 *__restrict__ b, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        b[i] = a[i];
}

__global__ void k_copy1(const uint8_t *__restrict__ a, uint8_t *__restrict__ b, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = a[i];
}

template <class T> __global__ void k_fill(T *__restrict__ p, uint64_t n, T v) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = v;
}

// Rows gathered by a segment list: output rows [seg_start[g], seg_start[g+1])
// are rows seg_row[g] .. of source relation seg_src[g] (all sources share one
// schema). A block takes kGatherChunk consecutive output rows: they span a
// contiguous range of segments, whose starts the block stages in LDS (at most
// one per row unless segments are empty; then it searches them in place), and
// each thread finds its rows' segments by binary search there. Every column then moves
// kGatherRpt rows per thread with all loads issued before the stores.
constexpr uint32_t kGatherThreads = 256, kGatherRpt = 8, kGatherChunk = 4096, kGatherMaxSrc = 16;

template <int R>
__device__ __forceinline__ void gather_col(const GatherSource *lsrc, uint32_t c, uint32_t t, void *dst,
                                           const uint32_t (&src)[R], const uint64_t (&row)[R],
                                           const uint64_t (&orow)[R], const bool (&on)[R]) {
    uint64_t lo[R], hi[R];
#pragma unroll
    for (int u = 0; u < R; ++u)
        load_elem(lsrc[src[u]].cols[c], t, row[u], lo[u], hi[u]);
#pragma unroll
    for (int u = 0; u < R; ++u)
        if (on[u])
            store_elem(dst, t, orow[u], lo[u], hi[u]);
}

// LSRC: the sources (at most kGatherMaxSrc) are staged in LDS; else read
// where they lie (one source per segment, as batched joins gather them).
template <bool LSRC>
__global__ __launch_bounds__(kGatherThreads) void k_gather_segments(const uint64_t *__restrict__ seg_start,
                                                                    uint32_t n_seg, const uint32_t *__restrict__ seg_src,
                                                                    const uint64_t *__restrict__ seg_row,
                                                                    const GatherSource *__restrict__ srcs, uint32_t n_src,
                                                                    uint32_t n_cols, uint64_t *__restrict__ out_keys,
                                                                    OutColSet out) {
    __shared__ uint64_t lstart[kGatherChunk];
    __shared__ GatherSource lsrc[kGatherMaxSrc];
    __shared__ uint32_t g_lo_s, n_here_s;
    const uint64_t total = seg_start[n_seg];
    const uint64_t c0 = static_cast<uint64_t>(blockIdx.x) * kGatherChunk;
    if (c0 >= total)
        return; // uniform
    const uint64_t c1 = min(total, c0 + kGatherChunk);
    if (threadIdx.x == 0) {
        const uint32_t a = upper_bound_idx(seg_start, n_seg, c0);
        const uint32_t b = upper_bound_idx(seg_start, n_seg, c1 - 1);
        g_lo_s = a;
        n_here_s = b - a + 1;
    }
    if (LSRC && threadIdx.x < n_src)
        lsrc[threadIdx.x] = srcs[threadIdx.x];
    __syncthreads();
    const GatherSource *sp = LSRC ? lsrc : srcs;
    const uint32_t g_lo = g_lo_s, n_here = n_here_s;
    // more segments than rows only with empty segments: then search in place
    const bool staged = n_here <= kGatherChunk;
    if (staged)
        for (uint32_t j = threadIdx.x; j < n_here; j += kGatherThreads)
            lstart[j] = seg_start[g_lo + j];
    __syncthreads();
    for (uint64_t base = c0; base < c1; base += kGatherThreads * kGatherRpt) {
        uint32_t src[kGatherRpt];
        uint64_t row[kGatherRpt], orow[kGatherRpt];
        bool on[kGatherRpt];
#pragma unroll
        for (int u = 0; u < static_cast<int>(kGatherRpt); ++u) {
            const uint64_t i = base + static_cast<uint64_t>(u) * kGatherThreads + threadIdx.x;
            on[u] = i < c1;
            const uint64_t x = on[u] ? i : c0;
            const uint32_t g = staged ? upper_bound_idx(lstart, n_here, x) : upper_bound_idx(seg_start + g_lo, n_here, x);
            src[u] = seg_src[g_lo + g];
            row[u] = seg_row[g_lo + g] + (x - seg_start[g_lo + g]);
            orow[u] = i;
        }
        {
            uint64_t kv[kGatherRpt];
#pragma unroll
            for (int u = 0; u < static_cast<int>(kGatherRpt); ++u)
                kv[u] = sp[src[u]].keys[row[u]];
#pragma unroll
            for (int u = 0; u < static_cast<int>(kGatherRpt); ++u)
                if (on[u])
                    out_keys[orow[u]] = kv[u];
        }
        for (uint32_t c = 0; c < n_cols; ++c)
            gather_col<kGatherRpt>(sp, c, col_type(out, c), col_ptr(out, c), src, row, orow, on);
    }
}

// Order-independent checksum: every output row's bytes (key | build cols |
// probe cols, src/Reference.cpp:22-43) are digested by two chained SplitMix64
// lanes; lane sums and the row count are accumulated mod 2^64.
__global__ __launch_bounds__(256) void k_checksum(uint64_t n, const uint64_t *__restrict__ keys, ColSet left, ColSet right,
                                                  uint32_t row_size, unsigned long long *__restrict__ acc) {
    __shared__ unsigned long long part[3][4];
    uint64_t s0 = 0, s1 = 0, cnt = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        Packer p(row_size);
        p.push_le8(keys[i], 8);
        for (uint32_t c = 0; c < left.n; ++c) {
            uint64_t lo, hi;
            const uint32_t t = col_type(left, c);
            load_elem(col_ptr(left, c), t, i, lo, hi);
            p.push(lo, hi, type_size(t));
        }
        for (uint32_t c = 0; c < right.n; ++c) {
            uint64_t lo, hi;
            const uint32_t t = col_type(right, c);
            load_elem(col_ptr(right, c), t, i, lo, hi);
            p.push(lo, hi, type_size(t));
        }
        p.finish();
        s0 += p.d.a;
        s1 += p.d.c;
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

void launch_hash_batch(hipStream_t s, const uint64_t *in, uint64_t n, uint64_t *out, KernelTimer *tm) {
    if (n == 0)
        return;
    TimedScope ts(tm, "hash_batch", s);
    const uint64_t pairs = (n + 1) / 2;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((pairs + 255) / 256, 256 * 64));
    k_hash_batch<<<grid, 256, 0, s>>>(in, n, out);
}

// Streaming rates of the part, for the roofline's measured ceiling: every
// 16-byte word of a buffer read once (nontemporal loads, eight in flight per
// thread; the words are folded into one value so the loads are kept), or
// written once.
typedef uint32_t v4u32 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_stream_read(const v4u32 *__restrict__ a, uint64_t n, uint32_t *__restrict__ sink) {
    constexpr int U = 8;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x * U;
    uint32_t acc = 0;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x * U + threadIdx.x; i < n; i += stride) {
        v4u32 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t j = i + static_cast<uint64_t>(u) * blockDim.x;
            v[u] = __builtin_nontemporal_load(a + (j < n ? j : 0));
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    if (acc == 0x9E3779B9u) // practically never; keeps the loads
        sink[0] = acc;
}

__global__ __launch_bounds__(256) void k_stream_write(v4u32 *__restrict__ b, uint64_t n) {
    constexpr int U = 4;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x * U;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x * U + threadIdx.x; i < n; i += stride) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t j = i + static_cast<uint64_t>(u) * blockDim.x;
            if (j < n)
                b[j] = v4u32{static_cast<uint32_t>(j), 1u, 2u, 3u};
        }
    }
}

__global__ __launch_bounds__(256) void k_stream_copy(const v4u32 *__restrict__ a, v4u32 *__restrict__ b, uint64_t n) {
    constexpr int U = 4;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x * U;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x * U + threadIdx.x; i < n; i += stride) {
        v4u32 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t j = i + static_cast<uint64_t>(u) * blockDim.x;
            v[u] = __builtin_nontemporal_load(a + (j < n ? j : 0));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t j = i + static_cast<uint64_t>(u) * blockDim.x;
            if (j < n)
                __builtin_nontemporal_store(v[u], b + j);
        }
    }
}

// ---- layout probes: the measurements behind the partitioning layout ----

// Random 16-byte reads of a table of `mask + 1` words (a power of two): read
// i goes to word int_hash64(seed + i) & mask, eight reads in flight per
// thread. The rate of lookups whose table lives in L2, the MALL or HBM.
__global__ __launch_bounds__(256) void k_probe_random_read(const v4u32 *__restrict__ t, uint64_t mask, uint64_t n,
                                                           uint64_t seed, uint32_t *__restrict__ sink) {
    constexpr int U = 8;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x * U;
    uint32_t acc = 0;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x * U + threadIdx.x; i < n; i += stride) {
        v4u32 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t j = i + static_cast<uint64_t>(u) * blockDim.x;
            v[u] = t[int_hash64(seed + j) & mask];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    if (acc == 0x9E3779B9u) // practically never; keeps the loads
        sink[0] = acc;
}

// Plain copy with cached loads and stores (a scratch buffer written and read
// back at once may stay in the caches).
__global__ __launch_bounds__(256) void k_probe_copy(const v4u32 *__restrict__ a, v4u32 *__restrict__ b, uint64_t n) {
    constexpr int U = 4;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x * U;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x * U + threadIdx.x; i < n; i += stride) {
        v4u32 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t j = i + static_cast<uint64_t>(u) * blockDim.x;
            v[u] = a[j < n ? j : 0];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t j = i + static_cast<uint64_t>(u) * blockDim.x;
            if (j < n)
                b[j] = v[u];
        }
    }
}

// The memory pattern of one radix pass over 32-byte rows (an 8-, an 8- and a
// 16-byte column), without the ranking: each 4,096-row tile sends its rows in
// runs of 4096 / parts consecutive rows to `parts` partitions of n / parts
// rows each (tile t's run of partition p at row p * n / parts + t * run), as
// the scatter's staged stores do. Input columns are read with nontemporal
// loads. n is a multiple of 4,096 and parts a power of two <= 4,096.
constexpr int kProbeTile = 4096;
constexpr int kProbeThreads = 1024;
__device__ __forceinline__ uint64_t probe_dest(uint64_t n, uint32_t parts, uint32_t tile, uint32_t k) {
    const uint32_t run = kProbeTile / parts;
    return static_cast<uint64_t>(k / run) * (n / parts) + static_cast<uint64_t>(tile) * run + k % run;
}
// Into three column outputs.
__global__ __launch_bounds__(kProbeThreads) void k_probe_pass_cols(const uint64_t *__restrict__ ik, const uint64_t *__restrict__ i1,
                                                                   const v4u32 *__restrict__ i2, uint64_t *__restrict__ ok,
                                                                   uint64_t *__restrict__ o1, v4u32 *__restrict__ o2, uint64_t n,
                                                                   uint32_t parts) {
    constexpr int R = kProbeTile / kProbeThreads;
    const uint32_t tile = blockIdx.x;
    const uint64_t base = static_cast<uint64_t>(tile) * kProbeTile;
    uint64_t a[R], b[R];
    v4u32 c[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        a[r] = __builtin_nontemporal_load(ik + base + r * kProbeThreads + threadIdx.x);
#pragma unroll
    for (int r = 0; r < R; ++r)
        ok[probe_dest(n, parts, tile, r * kProbeThreads + threadIdx.x)] = a[r];
    __syncthreads(); // a column at a time per workgroup, as the scatter's staged columns go
#pragma unroll
    for (int r = 0; r < R; ++r)
        b[r] = __builtin_nontemporal_load(i1 + base + r * kProbeThreads + threadIdx.x);
#pragma unroll
    for (int r = 0; r < R; ++r)
        o1[probe_dest(n, parts, tile, r * kProbeThreads + threadIdx.x)] = b[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r)
        c[r] = __builtin_nontemporal_load(i2 + base + r * kProbeThreads + threadIdx.x);
#pragma unroll
    for (int r = 0; r < R; ++r)
        o2[probe_dest(n, parts, tile, r * kProbeThreads + threadIdx.x)] = c[r];
}
// Into one row-major output: a lane pair writes a row's two 16-byte halves.
__global__ __launch_bounds__(kProbeThreads) void k_probe_pass_rows(const uint64_t *__restrict__ ik, const uint64_t *__restrict__ i1,
                                                                   const v4u32 *__restrict__ i2, v4u32 *__restrict__ orow,
                                                                   uint64_t n, uint32_t parts) {
    constexpr int R = 2 * kProbeTile / kProbeThreads;
    const uint32_t tile = blockIdx.x;
    const uint64_t base = static_cast<uint64_t>(tile) * kProbeTile;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t w = r * kProbeThreads + threadIdx.x;
        const uint32_t k = w >> 1;
        v4u32 v;
        if (w & 1u) {
            v = __builtin_nontemporal_load(i2 + base + k);
        } else {
            const uint64_t x = __builtin_nontemporal_load(ik + base + k);
            const uint64_t y = __builtin_nontemporal_load(i1 + base + k);
            v = v4u32{static_cast<uint32_t>(x), static_cast<uint32_t>(x >> 32), static_cast<uint32_t>(y),
                      static_cast<uint32_t>(y >> 32)};
        }
        orow[probe_dest(n, parts, tile, k) * 2 + (w & 1u)] = v;
    }
}

void launch_probe_random_read(hipStream_t s, const void *table, uint64_t words, uint64_t reads, uint64_t seed,
                              uint32_t *sink) {
    if (reads == 0 || words == 0)
        return;
    k_probe_random_read<<<256 * 16, 256, 0, s>>>(static_cast<const v4u32 *>(table), words - 1, reads, seed, sink);
}

void launch_probe_copy(hipStream_t s, void *dst, const void *src, uint64_t bytes) {
    const uint64_t n16 = bytes / 16;
    if (n16 == 0)
        return;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((n16 + 1023) / 1024, 256 * 64));
    k_probe_copy<<<grid, 256, 0, s>>>(static_cast<const v4u32 *>(src), static_cast<v4u32 *>(dst), n16);
}

void launch_probe_pass_cols(hipStream_t s, const void *in, uint64_t *ok, uint64_t *o1, void *o2, uint64_t n,
                            uint32_t parts) {
    if (n == 0)
        return;
    const uint32_t tiles = checked_grid(n / kProbeTile, kProbeThreads, "probe_pass");
    const uint64_t *ik = static_cast<const uint64_t *>(in);
    k_probe_pass_cols<<<tiles, kProbeThreads, 0, s>>>(ik, ik + n, reinterpret_cast<const v4u32 *>(ik + 2 * n), ok, o1,
                                                      static_cast<v4u32 *>(o2), n, parts);
}

void launch_probe_pass(hipStream_t s, bool rows, const void *in, void *out, uint64_t n, uint32_t parts) {
    if (n == 0)
        return;
    const uint32_t tiles = checked_grid(n / kProbeTile, kProbeThreads, "probe_pass");
    const uint64_t *ik = static_cast<const uint64_t *>(in);
    const uint64_t *i1 = ik + n;
    const v4u32 *i2 = reinterpret_cast<const v4u32 *>(i1 + n);
    if (rows) {
        k_probe_pass_rows<<<tiles, kProbeThreads, 0, s>>>(ik, i1, i2, static_cast<v4u32 *>(out), n, parts);
    } else {
        uint64_t *ok = static_cast<uint64_t *>(out);
        uint64_t *o1 = ok + n;
        k_probe_pass_cols<<<tiles, kProbeThreads, 0, s>>>(ik, i1, i2, ok, o1, reinterpret_cast<v4u32 *>(o1 + n), n, parts);
    }
}

void launch_stream(hipStream_t s, int mode, void *dst, const void *src, uint64_t bytes, KernelTimer *tm) {
    const uint64_t n16 = bytes / 16;
    if (n16 == 0)
        return;
    TimedScope ts(tm, mode == 0 ? "stream_copy" : mode == 1 ? "stream_read" : "stream_write", s);
    if (mode == 0) {
        // four 16-byte words in flight per thread, nontemporal both ways:
        // the fastest of the copy forms measured on this part
        k_stream_copy<<<256 * 64, 256, 0, s>>>(static_cast<const v4u32 *>(src), static_cast<v4u32 *>(dst), n16);
        return;
    }
    const uint32_t grid = 256 * 16;
    if (mode == 1)
        k_stream_read<<<grid, 256, 0, s>>>(static_cast<const v4u32 *>(src), n16, static_cast<uint32_t *>(dst));
    else
        k_stream_write<<<grid, 256, 0, s>>>(static_cast<v4u32 *>(dst), n16);
}

void launch_copy(hipStream_t s, void *dst, const void *src, uint64_t bytes, KernelTimer *tm) {
    TimedScope ts(tm, "copy", s);
    const uint64_t n16 = bytes / 16;
    if (n16 > 0) {
        const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((n16 + 255) / 256, 256 * 64));
        k_copy16<<<grid, 256, 0, s>>>(static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n16);
    }
    const uint64_t tail = bytes - n16 * 16;
    if (tail > 0)
        k_copy1<<<1, 64, 0, s>>>(static_cast<const uint8_t *>(src) + n16 * 16, static_cast<uint8_t *>(dst) + n16 * 16, tail);
}

void launch_fill_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t v) {
    if (n == 0)
        return;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((n + 255) / 256, 256 * 32));
    k_fill<uint32_t><<<grid, 256, 0, s>>>(p, n, v);
}

void launch_fill_u64(hipStream_t s, uint64_t *p, uint64_t n, uint64_t v) {
    if (n == 0)
        return;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((n + 255) / 256, 256 * 32));
    k_fill<uint64_t><<<grid, 256, 0, s>>>(p, n, v);
}

void gather_segments(hipStream_t s, const uint64_t *seg_start, uint32_t n_seg, const uint32_t *seg_src,
                     const uint64_t *seg_row, const GatherSource *srcs, uint32_t n_src, uint64_t total_rows,
                     uint64_t *out_keys, const OutColSet &out, KernelTimer *tm) {
    if (total_rows == 0 || n_seg == 0)
        return;
    TimedScope ts(tm, "gather_segments", s);
    const uint32_t grid = checked_grid((total_rows + kGatherChunk - 1) / kGatherChunk, kGatherThreads, "gather_segments");
    if (n_src <= kGatherMaxSrc)
        k_gather_segments<true><<<grid, kGatherThreads, 0, s>>>(seg_start, n_seg, seg_src, seg_row, srcs, n_src, out.n,
                                                                out_keys, out);
    else
        k_gather_segments<false><<<grid, kGatherThreads, 0, s>>>(seg_start, n_seg, seg_src, seg_row, srcs, n_src, out.n,
                                                                 out_keys, out);
}

void checksum_columns(hipStream_t s, uint64_t n, const uint64_t *keys, const ColSet &left, const ColSet &right,
                      unsigned long long *acc3, KernelTimer *tm) {
    uint32_t row_size = 8;
    for (uint32_t c = 0; c < left.n; ++c)
        row_size += type_size(left.types[c]);
    for (uint32_t c = 0; c < right.n; ++c)
        row_size += type_size(right.types[c]);
    k_fill<unsigned long long><<<1, 64, 0, s>>>(acc3, 3, 0ull);
    if (n == 0)
        return;
    TimedScope ts(tm, "checksum", s);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((n + 255) / 256, 256 * 16));
    k_checksum<<<grid, 256, 0, s>>>(n, keys, left, right, row_size, acc3);
}

} // namespace phjd
