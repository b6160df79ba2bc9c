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
        const uint32_t t = col_type(rc, c);
        load_elem(col_ptr(rc, c), t, r, lo, hi);
        store_elem(col_ptr(out.left, c), t, pos, lo, hi);
    }
    for (uint32_t c = 0; c < sc.n; ++c) {
        uint64_t lo, hi;
        const uint32_t t = col_type(sc, c);
        load_elem(col_ptr(sc, c), t, j, lo, hi);
        store_elem(col_ptr(out.right, c), t, pos, lo, hi);
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

constexpr uint16_t kNoNext = 0xFFFFu;   // end of a chain / probe row without a match
constexpr uint16_t kMore = 0x8000u;     // the referenced row has a successor in its chain
constexpr uint16_t kRowMask = 0x7FFFu;  // leaf-local build row (leaves hold < 32768 rows)

__global__ void k_leaf_items(const uint32_t *__restrict__ work_first, uint32_t n_leaves, uint32_t *__restrict__ item_leaf) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n_leaves)
        for (uint32_t w = work_first[g]; w < work_first[g + 1]; ++w)
            item_leaf[w] = g;
}

struct Item {
    uint32_t leaf, item, r0, nr, s0, s1;
};

__device__ __forceinline__ bool locate_item(const LeafJoinArgs &a, uint32_t w, Item &it) {
    if (w >= a.work_first[a.n_leaves])
        return false;
    it.leaf = a.item_leaf[w];
    it.item = w - a.work_first[it.leaf];
    it.r0 = a.r_off[it.leaf];
    it.nr = a.r_off[it.leaf + 1] - it.r0;
    it.s0 = a.s_off[it.leaf] + it.item * a.chunk_rows;
    it.s1 = min(a.s_off[it.leaf + 1], it.s0 + a.chunk_rows);
    return true;
}

// Builds the leaf's open-addressing table of 16-byte cells in LDS (64-bit CAS
// on the key, exchange on the chain head; src/HashTable.h:84-109) and the
// leaf-local chain successors next[] (src/BlockStore.h:86-89). Ends with a
// barrier. Returns true in every thread when some key was inserted twice.
template <int THREADS>
__device__ __forceinline__ void build_leaf_table(const LeafJoinArgs &a, const Item &it, uint32_t cap, LdsCell *cells,
                                                 uint16_t *next, uint32_t &side_head, uint32_t &dup) {
    for (uint32_t i = threadIdx.x; i < cap; i += THREADS) {
        LdsCell c;
        c.key = kEmptyKey;
        c.block_no = it.leaf;
        c.row_no = kInvalidRow;
        cells[i] = c;
    }
    if (threadIdx.x == 0) {
        side_head = kInvalidRow;
        dup = 0;
    }
    __syncthreads();

    bool seen = false;
    for (uint32_t i = threadIdx.x; i < it.nr; i += THREADS) {
        const uint64_t k = a.r_keys[it.r0 + i];
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
        seen |= prev != kInvalidRow;
        next[i] = prev == kInvalidRow ? kNoNext : static_cast<uint16_t>(prev);
    }
    if (seen)
        dup = 1u;
    __syncthreads();
}

// Head of key k's chain (leaf-local build row) or kInvalidRow.
__device__ __forceinline__ uint32_t lookup_leaf(const LdsCell *cells, uint32_t cap, uint32_t side_head, uint64_t k) {
    if (k == kEmptyKey)
        return side_head;
    uint32_t pos = home_slot(int_hash64(k), cap);
    while (true) {
        const LdsCell c = cells[pos];
        if (c.key == k)
            return c.row_no;
        if (c.key == kEmptyKey)
            return kInvalidRow;
        pos = (pos + 1 == cap) ? 0 : pos + 1;
    }
}

// Probe pass over one (leaf, probe chunk) work item: build the leaf's table,
// look up every probe row, and record per probe row the matched build row
// (heads, 16 bits with a "chain continues" flag) and the item's output rows.
// Chains are exported as next16 so the emit pass needs no table.
template <int THREADS, int U>
__global__ __launch_bounds__(THREADS) void k_leaf_probe(LeafJoinArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t tmp[THREADS / 64 + 1];
    __shared__ uint32_t side_head;
    __shared__ uint32_t dup;
    __shared__ uint32_t blk_chain;
    Item it;
    const uint32_t w = xcd_share(blockIdx.x, a.work_first[a.n_leaves]);
    if (!locate_item(a, w, it))
        return;
    const uint32_t cap = table_cap(it.nr, a.load_pct);
    LdsCell *cells = reinterpret_cast<LdsCell *>(smem);
    uint16_t *next = reinterpret_cast<uint16_t *>(smem + static_cast<size_t>(cap) * sizeof(LdsCell));
    if (threadIdx.x == 0)
        blk_chain = 0;
    build_leaf_table<THREADS>(a, it, cap, cells, next, side_head, dup);
    // Export the chains (first work item of the leaf only): successor plus
    // whether the successor continues.
    if (it.item == 0) {
        for (uint32_t i = threadIdx.x; i < it.nr; i += THREADS) {
            const uint16_t nx = next[i];
            a.next16[it.r0 + i] = nx == kNoNext ? kNoNext : static_cast<uint16_t>(nx | (next[nx] != kNoNext ? kMore : 0));
        }
    }

    uint32_t cnt = 0;
    bool saw_chain = false;
    for (uint32_t base = it.s0; base < it.s1; base += THREADS * U) {
        uint64_t key[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = base + u * THREADS + threadIdx.x;
            key[u] = j < it.s1 ? a.s_keys[j] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = base + u * THREADS + threadIdx.x;
            if (j >= it.s1)
                continue;
            const uint32_t hd = lookup_leaf(cells, cap, side_head, key[u]);
            uint16_t h16 = kNoNext;
            if (hd != kInvalidRow) {
                h16 = static_cast<uint16_t>(hd | (next[hd] != kNoNext ? kMore : 0));
                if (h16 & kMore)
                    saw_chain = true;
                for (uint32_t x = hd; x != kInvalidRow;) {
                    ++cnt;
                    const uint16_t nx = next[x];
                    x = nx == kNoNext ? kInvalidRow : nx;
                }
            }
            a.heads[j] = h16;
        }
    }
    if (saw_chain)
        blk_chain = 1u;
    uint32_t t;
    block_excl_scan<THREADS>(cnt, tmp, &t);
    if (threadIdx.x == 0)
        a.item_count[w] = t;
    if (threadIdx.x == 0 && blk_chain)
        atomicOr(a.chain_flag, 1u);
}

template <class T> __device__ __forceinline__ const T *cptr(const ColSet &s, uint32_t c) {
    return static_cast<const T *>(col_ptr(s, c));
}
template <class T> __device__ __forceinline__ T *optr(const OutColSet &s, uint32_t c) {
    return static_cast<T *>(col_ptr(s, c));
}

// Column-major gather for the single-match rows of one round: every slot's
// load is issued before any store, for one column at a time.
template <int U, class T>
__device__ __forceinline__ void gather_col(const T *__restrict__ src, T *__restrict__ dst, const uint32_t (&idx)[U],
                                           const uint64_t (&pos)[U], const bool (&on)[U]) {
    T v[U] = {};
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (on[u])
            v[u] = src[idx[u]];
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (on[u])
            dst[pos[u]] = v[u];
}

// 16-byte columns as two 8-byte halves (keeps the staging in registers).
template <int U>
__device__ __forceinline__ void gather_col128(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst,
                                              const uint32_t (&idx)[U], const uint64_t (&pos)[U], const bool (&on)[U]) {
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        uint64_t v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            v[u] = 0;
            if (on[u])
                v[u] = src[2 * static_cast<uint64_t>(idx[u]) + half];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (on[u])
                dst[2 * pos[u] + half] = v[u];
    }
}

template <int U>
__device__ __forceinline__ void gather_set(const ColSet &in, const OutColSet &out, const uint32_t (&idx)[U],
                                           const uint64_t (&pos)[U], const bool (&on)[U]) {
    for (uint32_t c = 0; c < in.n; ++c) {
        switch (col_type(in, c)) {
        case 0: gather_col<U>(cptr<uint8_t>(in, c), optr<uint8_t>(out, c), idx, pos, on); break;
        case 1: gather_col<U>(cptr<uint16_t>(in, c), optr<uint16_t>(out, c), idx, pos, on); break;
        case 2: gather_col<U>(cptr<uint32_t>(in, c), optr<uint32_t>(out, c), idx, pos, on); break;
        case 3: gather_col<U>(cptr<uint64_t>(in, c), optr<uint64_t>(out, c), idx, pos, on); break;
        default: gather_col128<U>(cptr<uint64_t>(in, c), optr<uint64_t>(out, c), idx, pos, on); break;
        }
    }
}

// Output layout of one round of THREADS*U probe rows: first every probe row
// with exactly one match, slot-major (all lanes' slot 0, then slot 1, ...), so
// each column store of a slot is one contiguous run; then the rows with
// several matches, thread by thread. Both emit kernels derive the same layout.
template <int THREADS, int U>
struct RoundLayout {
    uint32_t single_pos[U]; // round-relative position of this thread's slot u (single match)
    uint32_t chain_pos;     // round-relative position of this thread's first multi-match row
    uint32_t total;
};

template <int THREADS, int U>
__device__ __forceinline__ void round_layout(const uint16_t (&h)[U], const uint32_t (&chainlen)[U],
                                             uint32_t (*wave_cnt)[THREADS / 64], uint32_t *tmp,
                                             RoundLayout<THREADS, U> &L) {
    constexpr int W = THREADS / 64;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    uint32_t wpre[U];
    uint32_t chain_cnt = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const bool single = h[u] != kNoNext && !(h[u] & kMore);
        const uint64_t m = __ballot(single);
        wpre[u] = mbcnt(m);
        if (lane == 0)
            wave_cnt[u][wave] = static_cast<uint32_t>(__popcll(m));
        chain_cnt += chainlen[u];
    }
    uint32_t chain_total;
    const uint32_t chain_off = block_excl_scan<THREADS>(chain_cnt, tmp, &chain_total); // barriers publish wave_cnt
    uint32_t running = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint32_t c = wave_cnt[u][w];
            pre += static_cast<uint32_t>(w) < wave ? c : 0u;
            tot += c;
        }
        L.single_pos[u] = running + pre + wpre[u];
        running += tot;
    }
    L.chain_pos = running + chain_off;
    L.total = running + chain_total;
}

// Emit pass: a streaming kernel over the probe rows of each work item using
// the recorded heads; no table. Output row = probe key | build columns of the
// matched row | probe columns (src/JoinOps.h:77-125), at the item's offset.
// CHAINS == false writes the single-match rows column by column (every load of
// a column issued before its stores); CHAINS == true derives the same layout
// and writes only the multi-match rows, walking their chains.
template <int THREADS, int U, bool CHAINS>
__global__ __launch_bounds__(THREADS) void k_leaf_emit(LeafJoinArgs a) {
    __shared__ uint32_t tmp[THREADS / 64 + 1];
    __shared__ uint32_t wave_cnt[2][U][THREADS / 64];
    Item it;
    const uint32_t w = xcd_share(blockIdx.x, a.work_first[a.n_leaves]);
    if (!locate_item(a, w, it))
        return;
    uint64_t round_base = a.out_base + a.item_off[w];
    uint32_t parity = 0;
    for (uint32_t base = it.s0; base < it.s1; base += THREADS * U, parity ^= 1) {
        uint16_t h[U];
        uint32_t chainlen[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = base + u * THREADS + threadIdx.x;
            h[u] = j < it.s1 ? a.heads[j] : kNoNext;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            chainlen[u] = 0;
            if (h[u] != kNoNext && (h[u] & kMore)) {
                uint32_t c = 1;
                for (uint16_t x = a.next16[it.r0 + (h[u] & kRowMask)];; x = a.next16[it.r0 + (x & kRowMask)]) {
                    ++c;
                    if (!(x & kMore))
                        break;
                }
                chainlen[u] = c;
            }
        }
        RoundLayout<THREADS, U> L;
        round_layout<THREADS, U>(h, chainlen, wave_cnt[parity], tmp, L);
        if (L.total == 0)
            continue;
        if (!CHAINS) {
            uint32_t ridx[U], sidx[U];
            uint64_t opos[U];
            bool on[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                on[u] = h[u] != kNoNext && !(h[u] & kMore);
                sidx[u] = base + u * THREADS + threadIdx.x;
                ridx[u] = it.r0 + (h[u] & kRowMask);
                opos[u] = round_base + L.single_pos[u];
            }
            {
                uint64_t kv[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    kv[u] = 0;
                    if (on[u])
                        kv[u] = a.s_keys[sidx[u]];
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (on[u])
                        a.out.keys[opos[u]] = kv[u];
            }
            gather_set<U>(a.r_cols, a.out.left, ridx, opos, on);
            gather_set<U>(a.s_cols, a.out.right, sidx, opos, on);
        } else {
            uint64_t pos = round_base + L.chain_pos;
            for (int u = 0; u < U; ++u) {
                if (chainlen[u] == 0)
                    continue;
                const uint32_t j = base + u * THREADS + threadIdx.x;
                const uint64_t k = a.s_keys[j];
                uint16_t x = h[u];
                while (true) {
                    emit_row(a.out, pos++, k, a.r_cols, it.r0 + (x & kRowMask), a.s_cols, j);
                    if (!(x & kMore))
                        break;
                    x = a.next16[it.r0 + (x & kRowMask)];
                }
            }
        }
        round_base += L.total;
    }
}

// Cell index of key k in the leaf table, kSideSlot for the empty-key
// sentinel's side chain, or kInvalidRow.
constexpr uint32_t kSideSlot = 0xFFFEu;
constexpr int kFusedRowsPerThread = 32;
__device__ __forceinline__ uint32_t lookup_slot(const LdsCell *cells, uint32_t cap, uint32_t side_head, uint64_t k) {
    if (k == kEmptyKey)
        return side_head == kInvalidRow ? kInvalidRow : kSideSlot;
    uint32_t pos = home_slot(int_hash64(k), cap);
    while (true) {
        const uint64_t ck = cells[pos].key;
        if (ck == k)
            return pos;
        if (ck == kEmptyKey)
            return kInvalidRow;
        pos = (pos + 1 == cap) ? 0 : pos + 1;
    }
}

// Fused leaf join for build sides without repeated keys (every probe row has
// at most one match). One workgroup per (leaf, probe chunk of up to
// THREADS*K rows): the chunk's probe keys are loaded first, the leaf's LDS
// table is built meanwhile, every row is looked up (the matched cell index
// is kept in registers), the item's output rows are reserved with one atomic
// add on the sink cursor, and rounds of THREADS*U rows are written slot-major
// (the emit pass's single-match layout). The output key is the matched
// cell's key. Rows at or beyond the sink capacity are counted, not written.
// A leaf with a repeated key raises *chain_flag and writes nothing: the
// caller then joins with the two-phase path.
template <int THREADS, int K, int U>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_leaf_join(LeafJoinArgs a) {
    static_assert(K % U == 0 && K % 2 == 0, "rounds of U slots");
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t tmp[THREADS / 64 + 1];
    __shared__ uint32_t wave_cnt[2][U][THREADS / 64];
    __shared__ uint32_t side_head;
    __shared__ uint32_t dup;
    __shared__ unsigned long long item_base;
    Item it;
    const uint32_t w = xcd_share(blockIdx.x, a.work_first[a.n_leaves]);
    if (!locate_item(a, w, it))
        return;
    const uint32_t cap = table_cap(it.nr, a.load_pct);
    LdsCell *cells = reinterpret_cast<LdsCell *>(smem);
    uint16_t *next = reinterpret_cast<uint16_t *>(smem + static_cast<size_t>(cap) * sizeof(LdsCell));

    // The first KP probe keys are loaded before the table is built.
    constexpr int KP = K / 2;
    uint64_t key[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        const uint32_t j = it.s0 + k * THREADS + threadIdx.x;
        key[k] = j < it.s1 ? a.s_keys[j] : 0;
    }
    build_leaf_table<THREADS>(a, it, cap, cells, next, side_head, dup);
    if (dup) {
        if (threadIdx.x == 0)
            atomicOr(a.chain_flag, 1u);
        return;
    }
    // slot code per row, two per word: 0 = no match, else cell index + 1
    uint32_t code[K / 2];
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t j = it.s0 + k * THREADS + threadIdx.x;
        if (k >= KP && k % KP == 0) {
            // the next KP keys, loaded once the previous ones are consumed
#pragma unroll
            for (int q = 0; q < KP; ++q) {
                const uint32_t jq = j + q * THREADS;
                key[q] = jq < it.s1 ? a.s_keys[jq] : 0;
            }
        }
        const uint32_t sl = j < it.s1 ? lookup_slot(cells, cap, side_head, key[k % KP]) : kInvalidRow;
        const uint32_t c = sl == kInvalidRow ? 0u : sl + 1;
        cnt += c != 0;
        if (k % 2 == 0)
            code[k / 2] = c;
        else
            code[k / 2] |= c << 16;
    }
    uint32_t total;
    block_excl_scan<THREADS>(cnt, tmp, &total);
    if (total == 0)
        return;
    if (threadIdx.x == 0)
        item_base = atomicAdd(a.out.cursor, static_cast<unsigned long long>(total));
    __syncthreads();
    uint64_t rb = item_base;
    const uint64_t out_cap = a.out.cap;
#pragma unroll
    for (int r = 0; r < K / U; ++r) {
        uint16_t h[U];
        uint32_t none[U];
        uint32_t c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = r * U + u;
            c[u] = (code[k / 2] >> (16 * (k % 2))) & 0xFFFFu;
            h[u] = c[u] ? 0 : kNoNext;
            none[u] = 0;
        }
        RoundLayout<THREADS, U> L;
        round_layout<THREADS, U>(h, none, wave_cnt[r & 1], tmp, L);
        if (L.total == 0)
            continue;
        uint32_t ridx[U], sidx[U];
        uint64_t opos[U], okey[U];
        bool on[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            opos[u] = rb + L.single_pos[u];
            on[u] = c[u] != 0 && opos[u] < out_cap;
            sidx[u] = it.s0 + (r * U + u) * THREADS + threadIdx.x;
            uint32_t row = 0;
            okey[u] = 0;
            if (on[u]) {
                if (c[u] - 1 == kSideSlot) {
                    okey[u] = kEmptyKey;
                    row = side_head;
                } else {
                    const LdsCell cell = cells[c[u] - 1];
                    okey[u] = cell.key;
                    row = cell.row_no;
                }
            }
            ridx[u] = it.r0 + row;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (on[u])
                a.out.keys[opos[u]] = okey[u];
        gather_set<U>(a.r_cols, a.out.left, ridx, opos, on);
        gather_set<U>(a.s_cols, a.out.right, sidx, opos, on);
        rb += L.total;
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
    return static_cast<size_t>(table_cap(lds_max_rows, load_pct)) * sizeof(LdsCell) +
           ((static_cast<size_t>(lds_max_rows) * 2 + 15) & ~size_t{15});
}

uint32_t leaf_join_lds_max_rows(uint32_t lds_bytes, uint32_t load_pct) {
    uint32_t lo = 0, hi = 65535; // chains are 16-bit row indices
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (leaf_join_lds_bytes(mid, load_pct) <= lds_bytes)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

void leaf_work_plan(hipStream_t s, const LeafJoinArgs &a, uint32_t *work_first, uint32_t *scan_tmp, uint32_t *item_leaf,
                    KernelTimer *tm) {
    TimedScope ts(tm, "join_plan", s);
    k_leaf_work_count<<<(a.n_leaves + 1 + 255) / 256, 256, 0, s>>>(a.r_off, a.s_off, a.n_leaves, a.chunk_rows,
                                                                    a.lds_max_rows, work_first);
    scan_u32(s, work_first, a.n_leaves + 1, nullptr, 1, scan_tmp, nullptr, "join_plan");
    k_leaf_items<<<(a.n_leaves + 255) / 256, 256, 0, s>>>(work_first, a.n_leaves, item_leaf);
}

template <int T, int U> static void launch_probe(hipStream_t s, const LeafJoinArgs &a, uint64_t max_work, size_t lds) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_leaf_probe<T, U>), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(lds));
    k_leaf_probe<T, U><<<grid8(max_work), T, lds, s>>>(a);
}

void leaf_probe(hipStream_t s, const LeafJoinArgs &a, uint64_t max_work, uint32_t threads, KernelTimer *tm) {
    if (max_work == 0)
        return;
    const size_t lds = leaf_join_lds_bytes(a.lds_max_rows, a.load_pct);
    TimedScope ts(tm, "leaf_probe", s);
    if (threads >= 1024)
        launch_probe<1024, 8>(s, a, max_work, lds);
    else if (threads >= 512)
        launch_probe<512, 16>(s, a, max_work, lds);
    else
        launch_probe<256, 16>(s, a, max_work, lds);
}

uint32_t leaf_join_fused_chunk() { return 512 * kFusedRowsPerThread; }

void leaf_join_fused(hipStream_t s, const LeafJoinArgs &a, uint64_t max_work, KernelTimer *tm) {
    if (max_work == 0)
        return;
    const size_t lds = leaf_join_lds_bytes(a.lds_max_rows, a.load_pct);
    TimedScope ts(tm, "leaf_join", s);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_leaf_join<512, kFusedRowsPerThread, 4>), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(lds));
    k_leaf_join<512, kFusedRowsPerThread, 4><<<grid8(max_work), 512, lds, s>>>(a);
}

void leaf_emit(hipStream_t s, const LeafJoinArgs &a, uint64_t max_work, bool chains, KernelTimer *tm) {
    if (max_work == 0)
        return;
    {
        TimedScope ts(tm, "leaf_emit", s);
        k_leaf_emit<256, 8, false><<<grid8(max_work), 256, 0, s>>>(a);
    }
    if (chains) {
        TimedScope ts(tm, "leaf_emit_chains", s);
        k_leaf_emit<256, 8, true><<<grid8(max_work), 256, 0, s>>>(a);
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
