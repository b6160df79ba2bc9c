// k_join.hip — build, probe, chain expansion and projection gather on gfx950.
//
// Two table placements share one emit path:
//   * LDS leaf join (PHJ leaves that fit a workgroup's LDS): one workgroup per
//     (leaf, probe chunk) builds the leaf's open-addressing table of 16-byte
//     (key, RowRefCell) cells in LDS with 64-bit CAS on the key slot, then
//     probes the chunk.
//   * HBM tables (CHJ's 256 lock-free sub-tables, oversize leaves, the
//     standalone JoinHashTable entry point): same cell format and protocol in
//     global memory.
// Semantics follow JoinHashTable::insert/find (src/HashTable.h:84-127) with the
// BlockStore next-chain (src/BlockStore.h:86-89, src/JoinOps.h:222-227): a
// repeated key swaps the head and threads the previous head into next[]; the
// probe walks head -> next -> ... emitting one row per match
// (src/JoinOps.h:260-271). The output row is probe key, build columns via the
// matched build row, probe columns via the probe row (src/JoinOps.h:77-125).
#include "kernels.h"

namespace phjd {

namespace {

struct LdsCell {
    uint64_t key;
    uint32_t block_no;
    uint32_t row_no;
};
static_assert(sizeof(LdsCell) == 16, "16-byte cell");

__host__ __device__ __forceinline__ uint32_t table_cap(uint32_t rows, uint32_t load_pct) {
    uint64_t cap = (static_cast<uint64_t>(rows) * 100 + load_pct - 1) / load_pct;
    if (cap <= rows)
        cap = rows + 1;
    return static_cast<uint32_t>(cap);
}

__device__ __forceinline__ uint32_t home_slot(uint64_t h, uint32_t cap) {
    return static_cast<uint32_t>((static_cast<uint64_t>(static_cast<uint32_t>(h)) * cap) >> 32);
}

// Appends the columns of one output row at `pos`.
__device__ __forceinline__ void emit_row(const OutSink &out, uint64_t pos, uint64_t key, const ColSet &rc, uint64_t r,
                                         const ColSet &sc, uint64_t j) {
    if (pos >= out.cap)
        return;
    out.keys[pos] = key;
    for (uint32_t c = 0; c < rc.n; ++c) {
        uint64_t lo, hi;
        load_elem(rc.ptr[c], rc.types[c], r, lo, hi);
        store_elem(out.left.ptr[c], rc.types[c], pos, lo, hi);
    }
    for (uint32_t c = 0; c < sc.n; ++c) {
        uint64_t lo, hi;
        load_elem(sc.ptr[c], sc.types[c], j, lo, hi);
        store_elem(out.right.ptr[c], sc.types[c], pos, lo, hi);
    }
}

__global__ void k_leaf_work_count(const uint32_t *__restrict__ r_off, const uint32_t *__restrict__ s_off, uint32_t n_leaves,
                                  uint32_t chunk, uint32_t lds_max, uint32_t *__restrict__ work) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n_leaves) {
        const uint32_t nr = r_off[g + 1] - r_off[g];
        const uint32_t ns = s_off[g + 1] - s_off[g];
        work[g] = (nr > 0 && ns > 0 && nr <= lds_max) ? (ns + chunk - 1) / chunk : 0u;
    } else if (g == n_leaves) {
        work[g] = 0;
    }
}

template <int THREADS, int U>
__global__ __launch_bounds__(THREADS) void k_leaf_join_lds(LeafJoinArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t tmp[THREADS / 64 + 1];
    __shared__ uint32_t side_head;
    __shared__ unsigned long long out_base;

    const uint32_t total = a.work_first[a.n_leaves];
    const uint32_t w = xcd_remap(blockIdx.x, gridDim.x);
    if (w >= total)
        return;
    const uint32_t leaf = upper_bound_idx(a.work_first, a.n_leaves, w);
    const uint32_t item = w - a.work_first[leaf];
    const uint32_t r0 = a.r_off[leaf];
    const uint32_t nr = a.r_off[leaf + 1] - r0;
    const uint32_t s0 = a.s_off[leaf] + item * a.chunk_rows;
    const uint32_t s1 = min(a.s_off[leaf + 1], s0 + a.chunk_rows);
    const uint32_t cap = table_cap(nr, a.load_pct);

    LdsCell *cells = reinterpret_cast<LdsCell *>(smem);
    uint32_t *next = reinterpret_cast<uint32_t *>(smem + static_cast<size_t>(cap) * sizeof(LdsCell));

    for (uint32_t i = threadIdx.x; i < cap; i += THREADS) {
        LdsCell c;
        c.key = kEmptyKey;
        c.block_no = leaf;
        c.row_no = kInvalidRow;
        cells[i] = c;
    }
    if (threadIdx.x == 0)
        side_head = kInvalidRow;
    __syncthreads();

    // Build: claim the key slot with a 64-bit CAS, then swap the chain head.
    for (uint32_t i = threadIdx.x; i < nr; i += THREADS) {
        const uint64_t k = a.r_keys[r0 + i];
        uint32_t prev;
        if (k == kEmptyKey) {
            prev = atomicExch(&side_head, i);
        } else {
            uint32_t pos = home_slot(int_hash64(k), cap);
            while (true) {
                const unsigned long long old = atomicCAS(reinterpret_cast<unsigned long long *>(&cells[pos].key),
                                                         static_cast<unsigned long long>(kEmptyKey),
                                                         static_cast<unsigned long long>(k));
                if (old == kEmptyKey || old == k) {
                    prev = atomicExch(&cells[pos].row_no, i);
                    break;
                }
                pos = (pos + 1 == cap) ? 0 : pos + 1;
            }
        }
        next[i] = prev;
    }
    __syncthreads();

    // Probe in rounds of THREADS*U rows: find heads, count chain lengths,
    // reserve one output range per round, expand and gather.
    for (uint32_t base = s0; base < s1; base += THREADS * U) {
        uint64_t key[U];
        uint32_t head[U];
        uint32_t cnt = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = base + u * THREADS + threadIdx.x;
            head[u] = kInvalidRow;
            key[u] = 0;
            if (j < s1) {
                const uint64_t k = a.s_keys[j];
                key[u] = k;
                uint32_t hd = kInvalidRow;
                if (k == kEmptyKey) {
                    hd = side_head;
                } else {
                    uint32_t pos = home_slot(int_hash64(k), cap);
                    while (true) {
                        const LdsCell c = cells[pos];
                        if (c.key == k) {
                            hd = c.row_no;
                            break;
                        }
                        if (c.key == kEmptyKey)
                            break;
                        pos = (pos + 1 == cap) ? 0 : pos + 1;
                    }
                }
                head[u] = hd;
                for (uint32_t x = hd; x != kInvalidRow; x = next[x])
                    ++cnt;
            }
        }
        uint32_t round_total;
        const uint32_t off = block_excl_scan<THREADS>(cnt, tmp, &round_total);
        if (round_total == 0)
            continue;
        if (threadIdx.x == 0)
            out_base = atomicAdd(a.out.cursor, static_cast<unsigned long long>(round_total));
        __syncthreads();
        uint64_t pos = out_base + off;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = base + u * THREADS + threadIdx.x;
            for (uint32_t x = head[u]; x != kInvalidRow; x = next[x])
                emit_row(a.out, pos++, key[u], a.r_cols, r0 + x, a.s_cols, j);
        }
    }
}

// ---------------------------------------------------------------- HBM tables

struct HtArgs {
    GlobalHT ht;
    HtMode mode;
    HtRows rows;
    uint32_t n_items; // kHtLeaves: number of tables (oversize leaves)
};

__device__ __forceinline__ uint64_t ht_total_rows(const HtArgs &a) {
    return a.mode == kHtLeaves ? a.rows.first[a.n_items] : a.rows.rows;
}

// Maps thread item g to (row index, table). Leaf mode: contiguous leaf ranges.
__device__ __forceinline__ void ht_locate(const HtArgs &a, uint64_t g, uint64_t &row, uint32_t &table, uint32_t &leaf) {
    if (a.mode == kHtLeaves) {
        const uint32_t b = upper_bound_idx(a.rows.first, a.n_items, static_cast<uint32_t>(g));
        leaf = a.rows.leaf_of[b];
        row = a.rows.seg_off[leaf] + (static_cast<uint32_t>(g) - a.rows.first[b]);
        table = b;
    } else {
        row = g;
        table = 0;
        leaf = 0;
    }
}

__device__ __forceinline__ void ht_home(const HtArgs &a, uint64_t h, uint32_t &table, uint64_t &base, uint64_t &mask,
                                        uint64_t &pos) {
    if (a.mode == kHtChj)
        table = static_cast<uint32_t>(h) & (a.ht.n_tables - 1);
    base = a.ht.cell_off[table];
    mask = a.ht.cell_off[table + 1] - base - 1;
    pos = (a.mode == kHtChj ? (h >> 8) : h) & mask;
}

__device__ __forceinline__ uint64_t ref_pack(const HtArgs &a, uint64_t row, uint32_t leaf) {
    uint32_t bno, rno;
    if (a.mode == kHtLeaves) {
        bno = leaf;
        rno = static_cast<uint32_t>(row - a.ht.build_seg_off[leaf]);
    } else {
        bno = static_cast<uint32_t>(row / a.ht.block_rows);
        rno = static_cast<uint32_t>(row % a.ht.block_rows);
    }
    return static_cast<uint64_t>(bno) | (static_cast<uint64_t>(rno) << 32);
}

__device__ __forceinline__ uint32_t ref_row(const HtArgs &a, uint64_t ref) {
    const uint32_t bno = static_cast<uint32_t>(ref);
    const uint32_t rno = static_cast<uint32_t>(ref >> 32);
    if (bno == kInvalidRow)
        return kInvalidRow;
    if (a.mode == kHtLeaves)
        return a.ht.build_seg_off[bno] + rno;
    return bno * a.ht.block_rows + rno;
}

__global__ void k_ht_init(Cell16 *__restrict__ cells, uint64_t n, uint32_t *__restrict__ side_head, uint32_t n_tables) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        Cell16 c;
        c.key = kEmptyKey;
        c.block_no = kInvalidRow;
        c.row_no = 0;
        cells[i] = c;
    }
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g < n_tables)
        side_head[g] = kInvalidRow;
}

__global__ __launch_bounds__(256) void k_chj_counts(const uint64_t *__restrict__ keys, uint64_t n, uint32_t n_tables,
                                                    uint32_t *__restrict__ counts) {
    __shared__ uint32_t c[256];
    for (uint32_t i = threadIdx.x; i < n_tables; i += 256)
        c[i] = 0;
    __syncthreads();
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(&c[static_cast<uint32_t>(int_hash64(keys[i])) & (n_tables - 1)], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n_tables; i += 256)
        if (c[i])
            atomicAdd(&counts[i], c[i]);
}

__global__ __launch_bounds__(256) void k_ht_build(HtArgs a) {
    const uint64_t n = ht_total_rows(a);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < n; g += stride) {
        uint64_t row;
        uint32_t table, leaf;
        ht_locate(a, g, row, table, leaf);
        const uint64_t k = a.rows.keys[row];
        const uint64_t h = int_hash64(k);
        uint64_t base, mask, pos;
        ht_home(a, h, table, base, mask, pos);
        const uint64_t myref = ref_pack(a, row, leaf);
        uint64_t prev;
        if (k == kEmptyKey) {
            prev = atomicExch(&a.ht.side_head[table], static_cast<uint32_t>(row));
            a.ht.next[row] = static_cast<uint32_t>(prev);
            continue;
        }
        while (true) {
            Cell16 *c = &a.ht.cells[base + pos];
            const unsigned long long old = atomicCAS(reinterpret_cast<unsigned long long *>(&c->key),
                                                     static_cast<unsigned long long>(kEmptyKey),
                                                     static_cast<unsigned long long>(k));
            if (old == kEmptyKey || old == k) {
                prev = atomicExch(reinterpret_cast<unsigned long long *>(&c->block_no), static_cast<unsigned long long>(myref));
                break;
            }
            pos = (pos + 1) & mask;
        }
        a.ht.next[row] = ref_row(a, prev);
    }
}

__device__ __forceinline__ uint32_t ht_lookup(const HtArgs &a, uint64_t k, uint32_t table) {
    const uint64_t h = int_hash64(k);
    uint64_t base, mask, pos;
    ht_home(a, h, table, base, mask, pos);
    if (k == kEmptyKey)
        return a.ht.side_head[table];
    while (true) {
        const Cell16 c = a.ht.cells[base + pos];
        if (c.key == k)
            return ref_row(a, static_cast<uint64_t>(c.block_no) | (static_cast<uint64_t>(c.row_no) << 32));
        if (c.key == kEmptyKey)
            return kInvalidRow;
        pos = (pos + 1) & mask;
    }
}

template <int THREADS, int U>
__global__ __launch_bounds__(THREADS) void k_ht_probe_emit(HtArgs a, ColSet rc, ColSet sc, OutSink out) {
    __shared__ uint32_t tmp[THREADS / 64 + 1];
    __shared__ unsigned long long out_base;
    const uint64_t n = ht_total_rows(a);
    const uint64_t per_round = static_cast<uint64_t>(gridDim.x) * THREADS * U;
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * THREADS * U; base < n; base += per_round) {
        uint64_t key[U];
        uint64_t prow[U];
        uint32_t head[U];
        uint32_t cnt = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t g = base + static_cast<uint64_t>(u) * THREADS + threadIdx.x;
            head[u] = kInvalidRow;
            key[u] = 0;
            prow[u] = 0;
            if (g < n) {
                uint64_t row;
                uint32_t table, leaf;
                ht_locate(a, g, row, table, leaf);
                const uint64_t k = a.rows.keys[row];
                key[u] = k;
                prow[u] = row;
                const uint32_t hd = ht_lookup(a, k, table);
                head[u] = hd;
                for (uint32_t x = hd; x != kInvalidRow; x = a.ht.next[x])
                    ++cnt;
            }
        }
        uint32_t round_total;
        const uint32_t off = block_excl_scan<THREADS>(cnt, tmp, &round_total);
        if (round_total == 0)
            continue;
        if (threadIdx.x == 0)
            out_base = atomicAdd(out.cursor, static_cast<unsigned long long>(round_total));
        __syncthreads();
        uint64_t pos = out_base + off;
#pragma unroll
        for (int u = 0; u < U; ++u)
            for (uint32_t x = head[u]; x != kInvalidRow; x = a.ht.next[x])
                emit_row(out, pos++, key[u], rc, x, sc, prow[u]);
    }
}

__global__ __launch_bounds__(256) void k_ht_find(HtArgs a, const uint64_t *__restrict__ keys, uint64_t n,
                                                 uint32_t *__restrict__ heads) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        heads[i] = ht_lookup(a, keys[i], 0);
}

__global__ void k_max_leaf(const uint32_t *__restrict__ off, uint32_t n, uint32_t *__restrict__ out) {
    __shared__ uint32_t m;
    if (threadIdx.x == 0)
        m = 0;
    __syncthreads();
    uint32_t local = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        local = max(local, off[i + 1] - off[i]);
    atomicMax(&m, local);
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(out, m);
}

__global__ void k_big_leaves(const uint32_t *__restrict__ r_off, const uint32_t *__restrict__ s_off, uint32_t n,
                             uint32_t lds_max, uint32_t *__restrict__ leaf_of, uint32_t *__restrict__ count) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) {
        const uint32_t nr = r_off[g + 1] - r_off[g];
        const uint32_t ns = s_off[g + 1] - s_off[g];
        if (nr > lds_max && ns > 0)
            leaf_of[atomicAdd(count, 1u)] = g;
    }
}

} // namespace

void leaf_max_rows(hipStream_t s, const uint32_t *off, uint32_t n, uint32_t *out) {
    launch_fill_u32(s, out, 1, 0);
    if (n == 0)
        return;
    k_max_leaf<<<std::min<uint32_t>((n + 255) / 256, 1024), 256, 0, s>>>(off, n, out);
}

void leaf_big_list(hipStream_t s, const uint32_t *r_off, const uint32_t *s_off, uint32_t n, uint32_t lds_max,
                   uint32_t *leaf_of, uint32_t *count) {
    launch_fill_u32(s, count, 1, 0);
    if (n == 0)
        return;
    k_big_leaves<<<(n + 255) / 256, 256, 0, s>>>(r_off, s_off, n, lds_max, leaf_of, count);
}

uint64_t leaf_max_work(uint64_t probe_rows, uint32_t n_leaves, uint32_t chunk_rows) {
    return probe_rows / chunk_rows + n_leaves;
}

size_t leaf_join_lds_bytes(uint32_t lds_max_rows, uint32_t load_pct) {
    return static_cast<size_t>(table_cap(lds_max_rows, load_pct)) * sizeof(LdsCell) + static_cast<size_t>(lds_max_rows) * 4;
}

void leaf_work_plan(hipStream_t s, const LeafJoinArgs &a, uint32_t *work_first, uint32_t *scan_tmp, KernelTimer *tm) {
    {
        TimedScope ts(tm, "join_plan", s);
        k_leaf_work_count<<<(a.n_leaves + 1 + 255) / 256, 256, 0, s>>>(a.r_off, a.s_off, a.n_leaves, a.chunk_rows,
                                                                        a.lds_max_rows, work_first);
    }
    scan_u32(s, work_first, a.n_leaves + 1, nullptr, 1, scan_tmp, tm, "join_plan");
}

void leaf_join_lds(hipStream_t s, const LeafJoinArgs &a, uint64_t max_work, uint32_t threads, KernelTimer *tm) {
    if (max_work == 0)
        return;
    const size_t lds = leaf_join_lds_bytes(a.lds_max_rows, a.load_pct);
    TimedScope ts(tm, "leaf_join", s);
    if (threads >= 1024) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_leaf_join_lds<1024, 2>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        k_leaf_join_lds<1024, 2><<<grid8(max_work), 1024, lds, s>>>(a);
    } else if (threads >= 512) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_leaf_join_lds<512, 4>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        k_leaf_join_lds<512, 4><<<grid8(max_work), 512, lds, s>>>(a);
    } else {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_leaf_join_lds<256, 4>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        k_leaf_join_lds<256, 4><<<grid8(max_work), 256, lds, s>>>(a);
    }
}

void ht_chj_plan(hipStream_t s, const uint64_t *keys, uint64_t rows, uint32_t n_tables, uint32_t *counts, KernelTimer *tm) {
    launch_fill_u32(s, counts, n_tables, 0);
    if (rows == 0)
        return;
    TimedScope ts(tm, "chj_plan", s);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((rows + 255) / 256, 1024));
    k_chj_counts<<<grid, 256, 0, s>>>(keys, rows, n_tables, counts);
}

void ht_init(hipStream_t s, const GlobalHT &ht, uint64_t total_cells, uint64_t, KernelTimer *tm) {
    TimedScope ts(tm, "ht_init", s);
    const uint64_t n = std::max<uint64_t>(total_cells, ht.n_tables);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((n + 255) / 256, 256 * 64));
    k_ht_init<<<grid, 256, 0, s>>>(ht.cells, total_cells, ht.side_head, ht.n_tables);
}

static HtArgs make_args(const GlobalHT &ht, HtMode mode, const HtRows &r) {
    HtArgs a;
    a.ht = ht;
    a.mode = mode;
    a.rows = r;
    a.n_items = ht.n_tables;
    return a;
}

void ht_build(hipStream_t s, const GlobalHT &ht, HtMode mode, const HtRows &r, uint64_t max_rows, KernelTimer *tm,
              const char *name) {
    if (max_rows == 0)
        return;
    TimedScope ts(tm, name, s);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((max_rows + 255) / 256, 256 * 64));
    k_ht_build<<<grid, 256, 0, s>>>(make_args(ht, mode, r));
}

void ht_probe_emit(hipStream_t s, const GlobalHT &ht, HtMode mode, const HtRows &probe, uint64_t max_rows,
                   const ColSet &r_cols, const ColSet &s_cols, const OutSink &out, KernelTimer *tm, const char *name) {
    if (max_rows == 0)
        return;
    TimedScope ts(tm, name, s);
    constexpr int T = 256, U = 4;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((max_rows + T * U - 1) / (T * U), 256 * 16));
    k_ht_probe_emit<T, U><<<grid, T, 0, s>>>(make_args(ht, mode, probe), r_cols, s_cols, out);
}

void ht_find(hipStream_t s, const GlobalHT &ht, const uint64_t *keys, uint64_t n, Cell16 *, uint32_t *heads_rows,
             KernelTimer *tm) {
    if (n == 0)
        return;
    TimedScope ts(tm, "ht_find", s);
    HtRows r{};
    r.keys = keys;
    r.rows = n;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((n + 255) / 256, 256 * 64));
    k_ht_find<<<grid, 256, 0, s>>>(make_args(ht, kHtSingle, r), keys, n, heads_rows);
}

} // namespace phjd
