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
// schema). Each thread finds its row's segment by binary search; neighbouring
// rows of a segment are neighbours in both the source and the output.
__global__ __launch_bounds__(256) void k_gather_segments(const uint64_t *__restrict__ seg_start, uint32_t n_seg,
                                                         const uint32_t *__restrict__ seg_src,
                                                         const uint64_t *__restrict__ seg_row,
                                                         const GatherSource *__restrict__ srcs, uint32_t n_cols,
                                                         uint64_t *__restrict__ out_keys, OutColSet out) {
    const uint64_t total = seg_start[n_seg];
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint32_t g = upper_bound_idx(seg_start, n_seg, i);
        const GatherSource &src = srcs[seg_src[g]];
        const uint64_t r = seg_row[g] + (i - seg_start[g]);
        out_keys[i] = src.keys[r];
        for (uint32_t c = 0; c < n_cols; ++c) {
            uint64_t lo, hi;
            const uint32_t t = col_type(out, c);
            load_elem(src.cols[c], t, r, lo, hi);
            store_elem(col_ptr(out, c), t, i, lo, hi);
        }
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
                     const uint64_t *seg_row, const GatherSource *srcs, uint64_t total_rows, uint64_t *out_keys,
                     const OutColSet &out, KernelTimer *tm) {
    if (total_rows == 0 || n_seg == 0)
        return;
    TimedScope ts(tm, "gather_segments", s);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((total_rows + 255) / 256, 256 * 32));
    k_gather_segments<<<grid, 256, 0, s>>>(seg_start, n_seg, seg_src, seg_row, srcs, out.n, out_keys, out);
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
