// k_join.hip — build, probe, chain expansion and projection gather on gfx950.
//
// Three ways to join, one output row format:
//   * Fused leaf join (k_leaf_join; PHJ leaves that fit LDS, no repeated build
//     key): one workgroup per (leaf, probe chunk) builds a compact table of the
//     leaf in LDS (u64 keys + u16 rows, 64-bit CAS on the key slot), looks up
//     its probe rows and writes their output rows, all in one launch.
//   * Two-phase leaf join (k_leaf_probe + k_leaf_emit; repeated build keys):
//     the probe pass builds 16-byte (key, RowRefCell) cells with next chains in
//     LDS, counts each work item's output rows and exports the leaf's rows
//     grouped by key; the emit pass writes rows at exact scanned offsets.
//   * HBM tables (CHJ's 256 lock-free sub-tables, oversize leaves, the
//     standalone JoinHashTable entry point): 16-byte cells and next chains in
//     global memory, the same insert protocol.
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

// Home cell of key k in a leaf's LDS table of `cap` cells. The leaf's keys
// already share the top bits of intHash64, so the table mixes the key itself:
// the 32-bit fold of k times a Fibonacci constant, scaled to [0, cap). The
// cell layout of a leaf table is internal (only the rows a probe finds are
// observable), so it need not follow JoinHashTable's home rule.
__device__ __forceinline__ uint32_t home_slot(uint64_t k, uint32_t cap) {
    const uint32_t x = (static_cast<uint32_t>(k) ^ static_cast<uint32_t>(k >> 32)) * 0x9E3779B1u;
    return __umulhi(x, cap);
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

// Per leaf: chain-export slots of the two-phase join, one region of nr rows per
// work item of the leaf (each item builds its own table, so chain orders of
// two items of one leaf need not agree).
__global__ void k_leaf_chain_slots(const uint32_t *__restrict__ r_off, const uint32_t *__restrict__ work_first,
                                   uint32_t n_leaves, uint32_t *__restrict__ chain_first) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n_leaves)
        chain_first[g] = (work_first[g + 1] - work_first[g]) * (r_off[g + 1] - r_off[g]);
    else if (g == n_leaves)
        chain_first[g] = 0;
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
    // the leaf's probe rows split evenly over its items (each <= chunk_rows)
    const uint32_t sl0 = a.s_off[it.leaf];
    const uint64_t ns = a.s_off[it.leaf + 1] - sl0;
    const uint32_t items = a.work_first[it.leaf + 1] - a.work_first[it.leaf];
    it.s0 = sl0 + static_cast<uint32_t>(ns * it.item / items);
    it.s1 = sl0 + static_cast<uint32_t>(ns * (it.item + 1) / items);
    return true;
}

// u16 entries of an LDS array of n rows, padded to 16 bytes.
__host__ __device__ constexpr uint32_t round_up16(uint32_t n) { return (n + 7u) & ~7u; }

// First chain-export slot of a work item (two-phase join): the item exports
// 2 * nr entries, its groups (perm, nr) and their lengths (glen, nr).
__device__ __forceinline__ uint64_t item_chain_base(const LeafJoinArgs &a, const Item &it) {
    return 2 * (static_cast<uint64_t>(a.chain_first[it.leaf]) + static_cast<uint64_t>(it.item) * it.nr);
}

// Builds the leaf's open-addressing table of 16-byte cells in LDS (64-bit CAS
// on the key, exchange on the chain head; src/HashTable.h:84-109) and the
// leaf-local chain successors next[] (src/BlockStore.h:86-89). Ends with a
// barrier; `dup` (LDS) is then nonzero when some key was inserted twice.
template <int THREADS>
__device__ __forceinline__ void build_leaf_table(const LeafJoinArgs &a, const Item &it, uint32_t cap, LdsCell *cells,
                                                 uint16_t *next, uint32_t &side_head, uint32_t &dup) {
    // Build keys are loaded KB per thread at a time, all loads in flight
    // together; the first batch is issued before the table is cleared.
    constexpr int KB = 8;
    uint64_t kk[KB];
#pragma unroll
    for (int q = 0; q < KB; ++q) {
        const uint32_t i = q * THREADS + threadIdx.x;
        kk[q] = a.r_keys[it.r0 + min(i, it.nr - 1)]; // rows past the leaf are skipped below
    }
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
    for (uint32_t base = 0; base < it.nr; base += KB * THREADS) {
        if (base != 0) {
#pragma unroll
            for (int q = 0; q < KB; ++q) {
                const uint32_t i = base + q * THREADS + threadIdx.x;
                kk[q] = a.r_keys[it.r0 + min(i, it.nr - 1)];
            }
        }
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            const uint32_t i = base + q * THREADS + threadIdx.x;
            if (i >= it.nr)
                continue;
            const uint64_t k = kk[q];
            uint32_t prev;
            if (k == kEmptyKey) {
                prev = atomicExch(&side_head, i);
            } else {
                uint32_t pos = home_slot(k, cap);
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
    }
    if (seen)
        dup = 1u;
    __syncthreads();
}

// Head of key k's chain (leaf-local build row) or kInvalidRow.
__device__ __forceinline__ uint32_t lookup_leaf(const LdsCell *cells, uint32_t cap, uint32_t side_head, uint64_t k) {
    if (k == kEmptyKey)
        return side_head;
    uint32_t pos = home_slot(k, cap);
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
// look up every probe row, and record per probe row its match (heads, 16
// bits) and the item's output rows. Each key's chain (src/BlockStore.h:86-89,
// newest row first) is laid out once per item as one contiguous group of a
// row permutation: perm[start .. start + len) lists the chain in order, and
// glen[start] = len. A probe row with one match records the build row; one
// with several records start | kMore. The emit pass then reads groups instead
// of walking chains.
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
    uint16_t *gstart = next + round_up16(it.nr);  // per head row: its group's start
    uint16_t *glen_h = gstart + round_up16(it.nr); // per head row: its group's length
    if (threadIdx.x == 0)
        blk_chain = 0;
    build_leaf_table<THREADS>(a, it, cap, cells, next, side_head, dup);

    // Groups in cell order, then the empty-key sentinel's side chain (group
    // index cap). Each thread owns a contiguous range of group indices.
    uint16_t *perm = a.next16 + item_chain_base(a, it);
    uint16_t *glen = perm + it.nr;
    const uint32_t per = (cap + 1 + THREADS - 1) / THREADS;
    const uint32_t g0 = min(cap + 1, threadIdx.x * per), g1 = min(cap + 1, g0 + per);
    auto head_of = [&](uint32_t g) -> uint32_t { return g < cap ? cells[g].row_no : side_head; };
    uint32_t mine = 0;
    for (uint32_t g = g0; g < g1; ++g) {
        const uint32_t hd = head_of(g);
        if (hd == kInvalidRow)
            continue;
        uint32_t len = 1;
        for (uint16_t x = next[hd]; x != kNoNext; x = next[x])
            ++len;
        glen_h[hd] = static_cast<uint16_t>(len);
        mine += len;
    }
    uint32_t total;
    uint32_t start = block_excl_scan<THREADS>(mine, tmp, &total);
    for (uint32_t g = g0; g < g1; ++g) {
        const uint32_t hd = head_of(g);
        if (hd == kInvalidRow)
            continue;
        const uint32_t len = glen_h[hd];
        gstart[hd] = static_cast<uint16_t>(start);
        glen[start] = static_cast<uint16_t>(len);
        uint32_t e = start;
        for (uint32_t x = hd; x != kInvalidRow;) {
            perm[e++] = static_cast<uint16_t>(x);
            const uint16_t nx = next[x];
            x = nx == kNoNext ? kInvalidRow : nx;
        }
        start += len;
    }
    __syncthreads();

    uint32_t cnt = 0;
    bool saw_chain = false;
    for (uint32_t base = it.s0; base < it.s1; base += THREADS * U) {
        uint64_t key[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = base + u * THREADS + threadIdx.x;
            key[u] = a.s_keys[min(j, it.s1 - 1)];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = base + u * THREADS + threadIdx.x;
            if (j >= it.s1)
                continue;
            const uint32_t hd = lookup_leaf(cells, cap, side_head, key[u]);
            uint16_t h16 = kNoNext;
            if (hd != kInvalidRow) {
                if (next[hd] == kNoNext) {
                    h16 = static_cast<uint16_t>(hd);
                    ++cnt;
                } else {
                    h16 = static_cast<uint16_t>(gstart[hd] | kMore);
                    cnt += glen_h[hd];
                    saw_chain = true;
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
// load is issued before any store, for one column at a time. 16-byte columns
// move as one 16-byte access per row (as two 8-byte halves the compiler kept
// both halves' addresses live: 140 VGPRs instead of 88 for U = 8).
template <int U, class T>
__device__ __forceinline__ void gather_col(const T *__restrict__ src, T *__restrict__ dst, const uint32_t (&idx)[U],
                                           const uint64_t (&pos)[U], const bool (&on)[U]) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
        v[u] = src[on[u] ? idx[u] : 0u];
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (on[u])
            dst[pos[u]] = v[u];
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
        default: gather_col<U>(cptr<ulonglong2>(in, c), optr<ulonglong2>(out, c), idx, pos, on); break;
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
    uint32_t chain_start;   // round-relative position of the first multi-match row
    uint32_t chain_total;   // multi-match rows of the round
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
    L.chain_start = running;
    L.chain_total = chain_total;
    L.total = running + chain_total;
}

// Emit kernels' workgroup size and probe rows per thread per round (both
// emit kernels derive the same round layout from them). 512 x 4 keeps a
// round at 2,048 probe rows with eight waves per workgroup: with the round's
// probe rows staged in LDS (64 KB per workgroup on the skewed workload), two
// workgroups give a SIMD four waves, where 256 x 8 gave it two.
constexpr int kEmitThreads = 512;
constexpr int kEmitSlots = 4;
// Number of entries <= x among a[lo, lo + SPAN) (ascending). Each level reads
// FAN - 1 block ends together and keeps the block that holds the answer; the
// levels below FAN halve. Binary (FAN 2) by default: on the skewed workload's
// multi-match emission FAN 4, 8 and 16 (fewer dependent LDS rounds, more
// reads) measured within 1% of it.
template <uint32_t SPAN, uint32_t FAN = 2>
__device__ __forceinline__ uint32_t count_le(const uint32_t *a, uint32_t lo, uint32_t x) {
    if constexpr (SPAN <= 1) {
        return lo;
    } else if constexpr (SPAN < FAN) {
        return count_le<SPAN / 2, FAN>(a, lo + (a[lo + SPAN / 2 - 1] <= x ? SPAN / 2 : 0u), x);
    } else {
        constexpr uint32_t STEP = SPAN / FAN;
        uint32_t c = 0;
#pragma unroll
        for (uint32_t j = 1; j < FAN; ++j)
            c += a[lo + j * STEP - 1] <= x ? 1u : 0u;
        return count_le<STEP, FAN>(a, lo + c * STEP, x);
    }
}

// LDS for the staged build columns of the multi-match emit (beyond it they
// are gathered from HBM).
constexpr size_t kEmitRcolLds = 64 * 1024;
// LDS for a round's probe keys and leading probe columns (multi-match emit).
constexpr size_t kEmitScolLds = 32 * 1024;

// LDS of the emit kernels for the item's groups: perm and glen, nr each.
__host__ __device__ constexpr size_t emit_chain_lds(uint32_t lds_max_rows) {
    return 2 * static_cast<size_t>(round_up16(lds_max_rows)) * 2;
}

// Build columns of a leaf staged in LDS, one after another: column c takes
// lds_max_rows elements of its width, 16-byte aligned.
__host__ __device__ constexpr size_t emit_rcol_bytes(uint32_t lds_max_rows, uint32_t width) {
    return (static_cast<size_t>(lds_max_rows) * width + 15) & ~size_t{15};
}

// LDS of every column of a set staged for `rows` rows.
inline size_t emit_rcols_bytes(const ColSet &cs, uint32_t rows) {
    size_t b = 0;
    for (uint32_t c = 0; c < cs.n; ++c)
        b += emit_rcol_bytes(rows, type_size(cs.types[c]));
    return b;
}
__device__ __forceinline__ size_t emit_rcols_bytes_dev(const ColSet &cs, uint32_t rows) {
    size_t b = 0;
    for (uint32_t c = 0; c < cs.n; ++c)
        b += emit_rcol_bytes(rows, type_size(col_type(cs, c)));
    return b;
}

// Copies rows [r0, r0 + nr) of every build column to LDS (workgroup-wide).
template <int THREADS>
__device__ __forceinline__ void stage_rcols(const ColSet &rc, uint32_t r0, uint32_t nr, uint32_t lds_max_rows,
                                            unsigned char *dst) {
    size_t off = 0;
    for (uint32_t c = 0; c < rc.n; ++c) {
        const uint32_t t = col_type(rc, c);
        const void *src = col_ptr(rc, c);
        unsigned char *d = dst + off;
        off += emit_rcol_bytes(lds_max_rows, type_size(t));
        switch (t) {
        case 0:
            for (uint32_t i = threadIdx.x; i < nr; i += THREADS)
                reinterpret_cast<uint8_t *>(d)[i] = static_cast<const uint8_t *>(src)[r0 + i];
            break;
        case 1:
            for (uint32_t i = threadIdx.x; i < nr; i += THREADS)
                reinterpret_cast<uint16_t *>(d)[i] = static_cast<const uint16_t *>(src)[r0 + i];
            break;
        case 2:
            for (uint32_t i = threadIdx.x; i < nr; i += THREADS)
                reinterpret_cast<uint32_t *>(d)[i] = static_cast<const uint32_t *>(src)[r0 + i];
            break;
        case 3:
            for (uint32_t i = threadIdx.x; i < nr; i += THREADS)
                reinterpret_cast<uint64_t *>(d)[i] = static_cast<const uint64_t *>(src)[r0 + i];
            break;
        default:
            for (uint32_t i = threadIdx.x; i < 2 * nr; i += THREADS)
                reinterpret_cast<uint64_t *>(d)[i] = static_cast<const uint64_t *>(src)[2 * static_cast<uint64_t>(r0) + i];
            break;
        }
    }
}

// Copies rows [r0, r0 + n) of the probe keys and of the first ncols probe
// columns to LDS (workgroup-wide): the keys first (rows u64), then column c
// at emit_rcol_bytes(rows, width) steps.
template <int THREADS>
__device__ __forceinline__ void stage_scols(const uint64_t *keys, const ColSet &sc, uint32_t ncols, uint32_t r0, uint32_t n,
                                            uint32_t rows, unsigned char *dst) {
    for (uint32_t i = threadIdx.x; i < n; i += THREADS)
        reinterpret_cast<uint64_t *>(dst)[i] = keys[r0 + i];
    ColSet head = sc;
    head.n = ncols;
    stage_rcols<THREADS>(head, r0, n, rows, dst + emit_rcol_bytes(rows, 8));
}

// gather_set with the first `nlds` columns read from their LDS copy (idx
// relative to the copy) and the others from memory (idx + row0).
template <int U>
__device__ __forceinline__ void gather_set_lds(const unsigned char *base, uint32_t lds_max_rows, const ColSet &in,
                                               const OutColSet &out, const uint32_t (&idx)[U], const uint64_t (&pos)[U],
                                               const bool (&on)[U], uint32_t nlds = 0xFFFFFFFFu, uint32_t row0 = 0) {
    size_t off = 0;
    for (uint32_t c = 0; c < in.n; ++c) {
        if (c >= nlds) {
            uint32_t g[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                g[u] = idx[u] + row0;
            switch (col_type(in, c)) {
            case 0: gather_col<U>(cptr<uint8_t>(in, c), optr<uint8_t>(out, c), g, pos, on); break;
            case 1: gather_col<U>(cptr<uint16_t>(in, c), optr<uint16_t>(out, c), g, pos, on); break;
            case 2: gather_col<U>(cptr<uint32_t>(in, c), optr<uint32_t>(out, c), g, pos, on); break;
            case 3: gather_col<U>(cptr<uint64_t>(in, c), optr<uint64_t>(out, c), g, pos, on); break;
            default: gather_col<U>(cptr<ulonglong2>(in, c), optr<ulonglong2>(out, c), g, pos, on); break;
            }
            continue;
        }
        const unsigned char *src = base + off;
        off += emit_rcol_bytes(lds_max_rows, type_size(col_type(in, c)));
        switch (col_type(in, c)) {
        case 0: gather_col<U>(reinterpret_cast<const uint8_t *>(src), optr<uint8_t>(out, c), idx, pos, on); break;
        case 1: gather_col<U>(reinterpret_cast<const uint16_t *>(src), optr<uint16_t>(out, c), idx, pos, on); break;
        case 2: gather_col<U>(reinterpret_cast<const uint32_t *>(src), optr<uint32_t>(out, c), idx, pos, on); break;
        case 3: gather_col<U>(reinterpret_cast<const uint64_t *>(src), optr<uint64_t>(out, c), idx, pos, on); break;
        default: gather_col<U>(reinterpret_cast<const ulonglong2 *>(src), optr<ulonglong2>(out, c), idx, pos, on); break;
        }
    }
}

// Four consecutive output rows [p, p + 4) of one column. Packed: one store
// of the four values (4 bytes for u8 up to 4 x 16 bytes for u128), for a
// group whose rows are all written and whose first element is 4-aligned in a
// column whose base allows it; otherwise row by row. A wave's narrow
// row-per-lane stores (64 bytes per u8 instruction) cost about as much as its
// 8-byte ones, so packing makes a column's cost follow its bytes.
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint64_t u64x2_t __attribute__((ext_vector_type(2)));
template <class T>
__device__ __forceinline__ void store4(T *dst, uint64_t p, const T (&v)[4], const bool (&on)[4], bool packed) {
    if (packed) {
        if constexpr (sizeof(T) == 1) {
            const uint32_t x = static_cast<uint32_t>(v[0]) | static_cast<uint32_t>(v[1]) << 8 |
                               static_cast<uint32_t>(v[2]) << 16 | static_cast<uint32_t>(v[3]) << 24;
            *reinterpret_cast<uint32_t *>(dst + p) = x;
        } else if constexpr (sizeof(T) == 2) {
            const uint64_t x = static_cast<uint64_t>(v[0]) | static_cast<uint64_t>(v[1]) << 16 |
                               static_cast<uint64_t>(v[2]) << 32 | static_cast<uint64_t>(v[3]) << 48;
            *reinterpret_cast<uint64_t *>(dst + p) = x;
        } else if constexpr (sizeof(T) == 4) {
            *reinterpret_cast<u32x4_t *>(dst + p) = u32x4_t{v[0], v[1], v[2], v[3]};
        } else if constexpr (sizeof(T) == 8) {
            *reinterpret_cast<u64x2_t *>(dst + p) = u64x2_t{v[0], v[1]};
            *reinterpret_cast<u64x2_t *>(dst + p + 2) = u64x2_t{v[2], v[3]};
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r)
                dst[p + r] = v[r];
        }
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (on[r])
                dst[p + r] = v[r];
    }
}

// Column of four rows: values src[idx[r]] (LDS or memory) to rows [p, p + 4).
template <class T>
__device__ __forceinline__ void move4(const T *src, const uint32_t (&idx)[4], T *dst, uint64_t p, const bool (&on)[4],
                                      bool packed) {
    T v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
        v[r] = src[idx[r]];
    store4<T>(dst, p, v, on, packed);
}

// Every column of a set for four rows: the first `nlds` columns from their LDS
// copy at `base` (lds_rows elements per column, emit_rcol_bytes apart), the
// others from memory at idx + row0.
__device__ __forceinline__ void move4_set(const unsigned char *base, uint32_t lds_rows, uint32_t nlds, const ColSet &in,
                                          uint64_t row0, const OutColSet &out, const uint32_t (&idx)[4], uint64_t p,
                                          const bool (&on)[4], bool packed) {
    size_t off = 0;
    for (uint32_t c = 0; c < in.n; ++c) {
        const uint32_t t = col_type(in, c);
        if (c < nlds) { // LDS and memory sources in separate calls: one pointer for both would be a generic one
            const unsigned char *src = base + off;
            off += emit_rcol_bytes(lds_rows, type_size(t));
            switch (t) {
            case 0: move4<uint8_t>(reinterpret_cast<const uint8_t *>(src), idx, optr<uint8_t>(out, c), p, on, packed); break;
            case 1: move4<uint16_t>(reinterpret_cast<const uint16_t *>(src), idx, optr<uint16_t>(out, c), p, on, packed); break;
            case 2: move4<uint32_t>(reinterpret_cast<const uint32_t *>(src), idx, optr<uint32_t>(out, c), p, on, packed); break;
            case 3: move4<uint64_t>(reinterpret_cast<const uint64_t *>(src), idx, optr<uint64_t>(out, c), p, on, packed); break;
            default: move4<u64x2_t>(reinterpret_cast<const u64x2_t *>(src), idx, optr<u64x2_t>(out, c), p, on, packed); break;
            }
        } else {
            switch (t) {
            case 0: move4<uint8_t>(cptr<uint8_t>(in, c) + row0, idx, optr<uint8_t>(out, c), p, on, packed); break;
            case 1: move4<uint16_t>(cptr<uint16_t>(in, c) + row0, idx, optr<uint16_t>(out, c), p, on, packed); break;
            case 2: move4<uint32_t>(cptr<uint32_t>(in, c) + row0, idx, optr<uint32_t>(out, c), p, on, packed); break;
            case 3: move4<uint64_t>(cptr<uint64_t>(in, c) + row0, idx, optr<uint64_t>(out, c), p, on, packed); break;
            default: move4<u64x2_t>(cptr<u64x2_t>(in, c) + row0, idx, optr<u64x2_t>(out, c), p, on, packed); break;
            }
        }
    }
}

// Whether every column of `out` starts where 4-row groups of its elements are
// aligned for move4's packed stores (4 bytes for u8 up to 16 bytes).
__device__ __forceinline__ bool cols_pack_aligned(const OutColSet &out) {
    bool ok = true;
    for (uint32_t c = 0; c < out.n; ++c) {
        const uint32_t w = type_size(col_type(out, c));
        const uint32_t need = w * 4 < 16 ? w * 4 : 16;
        ok &= (reinterpret_cast<uint64_t>(col_ptr(out, c)) & (need - 1)) == 0;
    }
    return ok;
}

// Emit pass: a streaming kernel over the probe rows of each work item using
// the recorded heads; no table. Output row = probe key | build columns of the
// matched row | probe columns (src/JoinOps.h:77-125), at the item's offset.
// The item's groups (perm, glen) are first copied to LDS. CHAINS == false
// writes the single-match rows column by column (every load of a column
// issued before its stores); CHAINS == true derives the same layout and
// writes only the multi-match rows: each thread publishes where its groups
// land in the round's output, and every thread then takes consecutive output
// rows, finding each row's group by binary search. It reads the leaf's build
// columns and the round's probe keys and leading probe columns from LDS
// copies (when they fit), so its emission loop issues stores and no memory
// loads: on gfx9 one counter orders a wave's loads and stores, and a load
// there would wait for every store issued before it.
template <int THREADS, int U, bool CHAINS>
__global__ __launch_bounds__(THREADS) void k_leaf_emit(LeafJoinArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t tmp[THREADS / 64 + 1];
    __shared__ uint32_t wave_cnt[2][U][THREADS / 64];
    Item it;
    const uint32_t w = xcd_share(blockIdx.x, a.work_first[a.n_leaves]);
    if (!locate_item(a, w, it))
        return;
    uint16_t *perm = reinterpret_cast<uint16_t *>(smem);
    uint16_t *glen = perm + round_up16(a.lds_max_rows);
    uint32_t *seg_end = reinterpret_cast<uint32_t *>(smem + emit_chain_lds(a.lds_max_rows)); // [THREADS * U]
    uint16_t *seg_gs = reinterpret_cast<uint16_t *>(smem + emit_chain_lds(a.lds_max_rows) + THREADS * U * 4);
    uint16_t *seg_sx = reinterpret_cast<uint16_t *>(smem + emit_chain_lds(a.lds_max_rows) + THREADS * U * 6);
    unsigned char *rlds = smem + emit_chain_lds(a.lds_max_rows) + THREADS * U * 8; // staged build columns
    unsigned char *slds = rlds + (CHAINS && a.r_lds ? emit_rcols_bytes_dev(a.r_cols, a.lds_max_rows) : 0); // round's probe rows
    {
        const uint16_t *src = a.next16 + item_chain_base(a, it);
        for (uint32_t i = threadIdx.x; i < it.nr; i += THREADS) {
            perm[i] = src[i];
            glen[i] = src[it.nr + i];
        }
        if (CHAINS && a.r_lds)
            stage_rcols<THREADS>(a.r_cols, it.r0, it.nr, a.lds_max_rows, rlds);
        __syncthreads();
    }
    uint64_t round_base = a.out_base + a.item_off[w];
    uint32_t parity = 0;
    const bool pack_ok = CHAINS && (reinterpret_cast<uint64_t>(a.out.keys) & 15) == 0 && cols_pack_aligned(a.out.left) &&
                         cols_pack_aligned(a.out.right);
    for (uint32_t base = it.s0; base < it.s1; base += THREADS * U, parity ^= 1) {
        uint16_t h[U];
        uint32_t chainlen[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = base + u * THREADS + threadIdx.x;
            const uint16_t hv = a.heads[min(j, it.s1 - 1)];
            h[u] = j < it.s1 ? hv : kNoNext;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            chainlen[u] = (h[u] != kNoNext && (h[u] & kMore)) ? glen[h[u] & kRowMask] : 0u;
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
                for (int u = 0; u < U; ++u)
                    kv[u] = a.s_keys[on[u] ? sidx[u] : 0u];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (on[u])
                        a.out.keys[opos[u]] = kv[u];
            }
            gather_set<U>(a.r_cols, a.out.left, ridx, opos, on);
            gather_set<U>(a.s_cols, a.out.right, sidx, opos, on);
        } else if (L.chain_total > 0) {
            // The round's multi-match slots as a compact list of segments in
            // output order (thread by thread, slot by slot): segment s holds
            // the rows of one slot's group, seg_end[s] is its exclusive end
            // (round-relative), seg_gs its group start, seg_sx its probe row.
            // Output row o belongs to the first segment ending after o, and
            // since no listed segment is empty, row o + 1 lies in that segment
            // or the next one.
            uint32_t ne = 0;
#pragma unroll
            for (int u = 0; u < U; ++u)
                ne += chainlen[u] > 0 ? 1u : 0u;
            uint32_t n_seg;
            uint32_t k = block_excl_scan<THREADS>(ne, tmp, &n_seg);
            uint32_t end = L.chain_pos - L.chain_start;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (chainlen[u] > 0) {
                    end += chainlen[u];
                    seg_end[k] = end;
                    seg_gs[k] = h[u] & kRowMask;
                    seg_sx[k] = static_cast<uint16_t>(u * THREADS + threadIdx.x);
                    ++k;
                }
            }
            // The round's probe keys (and the leading probe columns that fit)
            // go to LDS with coalesced loads: the emission below then issues
            // no memory loads, only stores, so no load waits behind them.
            if (a.s_lds)
                stage_scols<THREADS>(a.s_keys, a.s_cols, a.s_lds - 1, base, min(static_cast<uint32_t>(THREADS * U), it.s1 - base),
                                     THREADS * U, slds);
            __syncthreads();
            // Output rows in aligned groups of four consecutive rows per
            // thread: the group's rows mostly share a segment (one search,
            // then a step to the next segment), and each column of a full
            // group leaves as one packed store.
            const uint64_t pos0 = round_base + L.chain_start;
            const uint32_t lead = static_cast<uint32_t>(pos0 & 3u); // rows of the first group before pos0
            const uint64_t a0 = pos0 - lead;
            const uint32_t ngroups = (lead + L.chain_total + 3) / 4;
            for (uint32_t g0 = 0; g0 < ngroups; g0 += THREADS) {
                const uint32_t g = g0 + threadIdx.x;
                const int64_t o0 = static_cast<int64_t>(g) * 4 - lead; // chain index of the group's first row
                uint32_t oc = static_cast<uint32_t>(min<int64_t>(max<int64_t>(o0, 0), L.chain_total - 1));
                // first segment ending after oc (the same number of steps in every lane)
                uint32_t sg = 0;
                for (uint32_t n = n_seg; n > 0;) {
                    const uint32_t half = n >> 1;
                    if (seg_end[sg + half] <= oc) {
                        sg += half + 1;
                        n -= half + 1;
                    } else {
                        n = half;
                    }
                }
                uint32_t ridx[4], sidx[4];
                bool on[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t o = o0 + r;
                    on[r] = g < ngroups && o >= 0 && o < static_cast<int64_t>(L.chain_total);
                    if (on[r])
                        oc = static_cast<uint32_t>(o); // rows without output repeat the last row with one
                    sg += oc >= seg_end[sg] ? 1u : 0u; // the row starts the next segment
                    const uint32_t e = oc - (sg ? seg_end[sg - 1] : 0u);
                    sidx[r] = seg_sx[sg];           // round-relative
                    ridx[r] = perm[seg_gs[sg] + e]; // leaf-local
                }
                const uint64_t p = a0 + static_cast<uint64_t>(g) * 4;
                const bool packed = pack_ok && on[0] && on[1] && on[2] && on[3];
                if (a.s_lds)
                    move4<uint64_t>(reinterpret_cast<const uint64_t *>(slds), sidx, a.out.keys, p, on, packed);
                else
                    move4<uint64_t>(a.s_keys + base, sidx, a.out.keys, p, on, packed);
                move4_set(rlds, a.lds_max_rows, a.r_lds ? a.r_cols.n : 0u, a.r_cols, it.r0, a.out.left, ridx, p, on, packed);
                move4_set(slds + emit_rcol_bytes(THREADS * U, 8), THREADS * U, a.s_lds ? a.s_lds - 1 : 0u, a.s_cols, base,
                          a.out.right, sidx, p, on, packed);
            }
            __syncthreads(); // the segment list is rewritten by the next round
        }
        round_base += L.total;
    }
}

// Fused leaf join: a matched slot of the leaf table, or kSideSlot for the
// empty-key sentinel's side chain (rows whose key is kEmptyKey).
constexpr uint32_t kSideSlot = 0xFFFEu;
// Probe rows per lane and super-rounds per work item (chunk = 512 threads).
constexpr int kFusedRowsPerThread = 16;
constexpr int kFusedSuperRounds = 1;
constexpr uint32_t kFusedChunkRows = kFusedSuperRounds * 512u * kFusedRowsPerThread;

// Element at byte offset `off` of a uniform base: the 32-bit offset keeps the
// address arithmetic in the scalar base + vector offset form.
template <class T> __device__ __forceinline__ T ld_at(const void *base, uint32_t off) {
    return *reinterpret_cast<const T *>(static_cast<const char *>(base) + off);
}
template <class T> __device__ __forceinline__ void st_at(void *base, uint32_t off, T v) {
    *reinterpret_cast<T *>(static_cast<char *>(base) + off) = v;
}

// One column of one round: rows idx[u] of src (element indices relative to
// src) to rows pos[u] of dst (relative to dst); loads first, then stores.
template <int U, class T>
__device__ __forceinline__ void move_col(const void *src, void *dst, const uint32_t (&idx)[U],
                                         const uint32_t (&pos)[U], const bool (&on)[U]) {
    // Unconditional loads (row 0 stands in for lanes without a row): a load
    // under a branch makes the compiler wait for it before the branch joins.
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
        v[u] = ld_at<T>(src, (on[u] ? idx[u] : 0u) * static_cast<uint32_t>(sizeof(T)));
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (on[u])
            st_at<T>(dst, pos[u] * static_cast<uint32_t>(sizeof(T)), v[u]);
}

// Every column of a set: source rows from `src_row0`, output rows from
// `dst_row0` (both uniform), per-lane indices relative to them.
template <int U>
__device__ __forceinline__ void move_set(const ColSet &in, uint64_t src_row0, const OutColSet &out, uint64_t dst_row0,
                                         const uint32_t (&idx)[U], const uint32_t (&pos)[U], const bool (&on)[U]) {
    for (uint32_t c = 0; c < in.n; ++c) {
        const uint32_t t = col_type(in, c);
        const uint64_t sz = type_size(t);
        const void *src = static_cast<const char *>(col_ptr(in, c)) + src_row0 * sz;
        void *dst = static_cast<char *>(col_ptr(out, c)) + dst_row0 * sz;
        switch (t) {
        case 0: move_col<U, uint8_t>(src, dst, idx, pos, on); break;
        case 1: move_col<U, uint16_t>(src, dst, idx, pos, on); break;
        case 2: move_col<U, uint32_t>(src, dst, idx, pos, on); break;
        case 3: move_col<U, uint64_t>(src, dst, idx, pos, on); break;
        default: move_col<U, ulonglong2>(src, dst, idx, pos, on); break; // one 16-byte access per row
        }
    }
}

// Per-slot emit metadata of the fused and prebuilt joins: bit 31 = the slot
// has an output row, bits 16..30 = its output row relative to the round's
// base, bits 0..15 = the matched build row within the leaf (or, before the
// fused join resolves it, the matched table slot).
__device__ __forceinline__ bool meta_on(uint32_t m) { return (m >> 31) != 0; }
__device__ __forceinline__ uint32_t meta_pos(uint32_t m) { return (m >> 16) & 0x7FFFu; }
__device__ __forceinline__ uint32_t meta_row(uint32_t m) { return m & 0xFFFFu; }

// Where a slot's source row comes from: the lane's probe row of the round
// (slot k is row k * THREADS + threadIdx.x), the build row in the metadata,
// or a separate array of (64-bit indexed) rows.
enum SweepRows { kRowProbe, kRowMeta, kRowArray };

template <int THREADS, int K, int MODE, class T, int G>
__device__ __forceinline__ void sweep_load(const void *src, const uint32_t (&meta)[K], const uint32_t (&rows)[K],
                                           uint32_t tid, int h, T (&v)[G]) {
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int k = h * G + u;
        if constexpr (MODE == kRowArray) {
            v[u] = static_cast<const T *>(src)[meta_on(meta[k]) ? rows[k] : 0u];
        } else {
            const uint32_t row = MODE == kRowProbe ? static_cast<uint32_t>(k * THREADS) + tid : meta_row(meta[k]);
            v[u] = ld_at<T>(src, (meta_on(meta[k]) ? row : 0u) * static_cast<uint32_t>(sizeof(T)));
        }
    }
}
template <int K, class T, int G>
__device__ __forceinline__ void sweep_store(void *dst, const uint32_t (&meta)[K], int h, const T (&v)[G]) {
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const int k = h * G + u;
        if (meta_on(meta[k]))
            st_at<T>(dst, meta_pos(meta[k]) * static_cast<uint32_t>(sizeof(T)), v[u]);
    }
}
// One column over all K slots of a lane, G slots at a time: the G loads are
// issued together, then the G stores, and group h + 1's loads go out before
// group h's stores (a row at a time would wait for each load before the next
// is issued, since the compiler cannot move a load above a store that may
// alias it). Slots without a row load row 0 and store nothing.
template <int THREADS, int K, int MODE, class T>
__device__ __forceinline__ void sweep_col(const void *src, void *dst, const uint32_t (&meta)[K],
                                          const uint32_t (&rows)[K], uint32_t tid) {
    constexpr int G0 = sizeof(T) >= 8 ? static_cast<int>(64 / sizeof(T)) : 16; // <= 16 VGPRs of values per group
    constexpr int G = G0 < K ? G0 : K;
    static_assert(K % G == 0, "slots in groups");
    // group h + 1's loads are issued before group h's stores
    T cur[G];
    sweep_load<THREADS, K, MODE, T, G>(src, meta, rows, tid, 0, cur);
#pragma unroll
    for (int h = 0; h < K / G; ++h) {
        T nxt[G];
        if (h + 1 < K / G)
            sweep_load<THREADS, K, MODE, T, G>(src, meta, rows, tid, h + 1, nxt);
        sweep_store<K, T, G>(dst, meta, h, cur);
        if (h + 1 < K / G) {
#pragma unroll
            for (int u = 0; u < G; ++u)
                cur[u] = nxt[u];
        }
    }
}

// Every column of a set (source rows from src_row0, output rows from dst_row0).
template <int THREADS, int K, int MODE>
__device__ __forceinline__ void sweep_set(const ColSet &in, uint64_t src_row0, const OutColSet &out, uint64_t dst_row0,
                                          const uint32_t (&meta)[K], const uint32_t (&rows)[K]) {
    for (uint32_t c = 0; c < in.n; ++c) {
        const uint32_t t = col_type(in, c);
        const uint64_t sz = type_size(t);
        const void *src = static_cast<const char *>(col_ptr(in, c)) + src_row0 * sz;
        void *dst = static_cast<char *>(col_ptr(out, c)) + dst_row0 * sz;
        // An opaque copy per column: offsets derived from the metadata are
        // then computed inside the column loop rather than hoisted out of it
        // for every column type at once (which spills).
        uint32_t m[K], r[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            m[k] = meta[k];
            asm volatile("" : "+v"(m[k]));
            r[k] = MODE == kRowArray ? rows[k] : 0u;
            if constexpr (MODE == kRowArray)
                asm volatile("" : "+v"(r[k]));
        }
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        switch (t) {
        case 0: sweep_col<THREADS, K, MODE, uint8_t>(src, dst, m, r, tid); break;
        case 1: sweep_col<THREADS, K, MODE, uint16_t>(src, dst, m, r, tid); break;
        case 2: sweep_col<THREADS, K, MODE, uint32_t>(src, dst, m, r, tid); break;
        case 3: sweep_col<THREADS, K, MODE, uint64_t>(src, dst, m, r, tid); break;
        default: sweep_col<THREADS, K, MODE, ulonglong2>(src, dst, m, r, tid); break;
        }
    }
}

// Leaf table of the fused join: keys[cap] (u64, kEmptyKey = free) and the
// leaf-local row of each key, rows[cap] (u16); no chains, since the fused
// join only runs without repeated keys. About 10 bytes per slot, so three
// 512-thread workgroups share a CU's LDS for leaves of ~2,500 rows.
__host__ __device__ __forceinline__ size_t fused_lds_bytes(uint32_t max_rows, uint32_t load_pct) {
    const size_t cap = table_cap(max_rows, load_pct);
    return cap * 8 + ((cap * 2 + 15) & ~size_t{15});
}
// The same rounded up to 16 bytes: the unit of a prebuilt table's copies.
__host__ __device__ __forceinline__ size_t fused_table_bytes(uint32_t rows, uint32_t load_pct) {
    return (fused_lds_bytes(rows, load_pct) + 15) & ~size_t{15};
}

// Builds the fused table (64-bit CAS on the key slot, then the row). Sets
// `dup` (LDS) when a key occurs twice. Ends with a barrier.
template <int THREADS>
__device__ __forceinline__ void build_fused_table(const LeafJoinArgs &a, const Item &it, uint32_t cap, uint64_t *tkeys,
                                                  uint16_t *trows, uint32_t &side_head, uint32_t &dup) {
    constexpr int KB = 8;
    uint64_t kk[KB];
#pragma unroll
    for (int q = 0; q < KB; ++q) {
        const uint32_t i = q * THREADS + threadIdx.x;
        kk[q] = a.r_keys[it.r0 + min(i, it.nr - 1)]; // rows past the leaf are skipped below
    }
    for (uint32_t i = threadIdx.x; i < cap; i += THREADS)
        tkeys[i] = kEmptyKey;
    if (threadIdx.x == 0) {
        side_head = kInvalidRow;
        dup = 0;
    }
    __syncthreads();
    bool seen = false;
    for (uint32_t base = 0; base < it.nr; base += KB * THREADS) {
        if (base != 0) {
#pragma unroll
            for (int q = 0; q < KB; ++q) {
                const uint32_t i = base + q * THREADS + threadIdx.x;
                kk[q] = a.r_keys[it.r0 + min(i, it.nr - 1)];
            }
        }
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            const uint32_t i = base + q * THREADS + threadIdx.x;
            if (i >= it.nr)
                continue;
            const uint64_t k = kk[q];
            if (k == kEmptyKey) {
                seen |= atomicExch(&side_head, i) != kInvalidRow;
                continue;
            }
            uint32_t pos = home_slot(k, cap);
            while (true) {
                const unsigned long long old = atomicCAS(reinterpret_cast<unsigned long long *>(&tkeys[pos]),
                                                         static_cast<unsigned long long>(kEmptyKey),
                                                         static_cast<unsigned long long>(k));
                if (old == kEmptyKey) {
                    trows[pos] = static_cast<uint16_t>(i);
                    break;
                }
                if (old == k) {
                    seen = true;
                    break;
                }
                pos = (pos + 1 == cap) ? 0 : pos + 1;
            }
        }
    }
    if (seen)
        dup = 1u;
    __syncthreads();
}

// Fused leaf join for build sides without repeated keys (every probe row has
// at most one match). One workgroup per (leaf, probe chunk) builds the leaf's
// compact LDS table, then walks the chunk in super-rounds of THREADS*K rows:
// every row is looked up (LB keys advance in lock step, one probe step per
// iteration, reads in flight together; the matched slot is kept in
// registers), the super-round's output rows are reserved with one atomic add
// on the sink cursor and laid out slot-major (slot k of every lane, wave by
// wave, so each column store of a slot is one contiguous run), then written
// in rounds of U slots. The output key is the matched slot's key. Rows at or
// beyond the sink capacity are counted, not written. A leaf with a repeated
// key raises the flag and writes nothing: the caller then joins with the
// two-phase path.
// Every column of a set gathered by absolute source rows (64-bit offsets):
// rows row[u] of each source column to rows dst_row0 + pos[u] of the output.
template <int U>
__device__ __forceinline__ void gather_set_abs(const ColSet &in, const OutColSet &out, uint64_t dst_row0,
                                               const uint32_t (&row)[U], const uint32_t (&pos)[U], const bool (&on)[U]) {
    for (uint32_t c = 0; c < in.n; ++c) {
        const uint32_t t = col_type(in, c);
        const void *src = col_ptr(in, c);
        void *dst = static_cast<char *>(col_ptr(out, c)) + dst_row0 * type_size(t);
        uint64_t lo[U], hi[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            load_elem(src, t, on[u] ? row[u] : 0u, lo[u], hi[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (on[u])
                store_elem(dst, t, pos[u], lo[u], hi[u]);
    }
}

// NARROW: the partitioned probe side carries only keys and each row's index
// in the unpartitioned probe relation (a.s_rowid); the probe columns of a
// matched row are gathered from a.s_src at that index.
template <int THREADS, int K, int U, bool NARROW>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(6))) void k_leaf_join(LeafJoinArgs a) {
    static_assert(K % U == 0 && K % 2 == 0 && K <= 64, "K slots per lane in rounds of U");
    constexpr int W = THREADS / 64;
    constexpr int LB = 4; // keys looked up in lock step
    static_assert(K % LB == 0, "lookup batches");
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t slot_off[K][W];
    __shared__ uint32_t side_head;
    __shared__ uint32_t dup;
    __shared__ uint32_t sr_total;
    __shared__ unsigned long long sr_base;
    Item it;
    const uint32_t w = xcd_share(blockIdx.x, a.work_first[a.n_leaves]);
    if (!locate_item(a, w, it))
        return;
    const uint32_t cap = table_cap(it.nr, a.load_pct);
    uint64_t *tkeys = reinterpret_cast<uint64_t *>(smem);
    uint16_t *trows = reinterpret_cast<uint16_t *>(smem + static_cast<size_t>(cap) * 8);
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;

    unsigned long long *cursor = a.out.cursor;
    uint64_t out_cap = a.out.cap;
    uint32_t *flag = a.chain_flag;
    if (a.leaf_join) {
        const uint32_t j = a.leaf_join[it.leaf];
        cursor = a.join_cursor + j;
        out_cap = a.join_end[j];
        flag = a.join_flag + j;
    }
    build_fused_table<THREADS>(a, it, cap, tkeys, trows, side_head, dup);
    if (dup) {
        if (threadIdx.x == 0)
            atomicOr(flag, 1u);
        return;
    }
#pragma unroll
    for (int sr = 0; sr < kFusedSuperRounds; ++sr) {
        const uint32_t sb = it.s0 + sr * THREADS * K;
        if (sb >= it.s1)
            break; // uniform
        // slot code per row, two per word: 0 = no match, else slot index + 1
        uint32_t code[K / 2];
        uint64_t key[LB], nkey[LB];
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            const uint32_t j = sb + q * THREADS + threadIdx.x;
            key[q] = a.s_keys[min(j, it.s1 - 1)]; // rows past the chunk are masked below
        }
#pragma unroll
        for (int g = 0; g < K / LB; ++g) {
            if (g + 1 < K / LB) {
#pragma unroll
                for (int q = 0; q < LB; ++q) {
                    const uint32_t j = sb + ((g + 1) * LB + q) * THREADS + threadIdx.x;
                    nkey[q] = a.s_keys[min(j, it.s1 - 1)];
                }
            }
            uint32_t pos[LB], res[LB];
            uint32_t pend = 0;
#pragma unroll
            for (int q = 0; q < LB; ++q) {
                const uint32_t j = sb + (g * LB + q) * THREADS + threadIdx.x;
                pos[q] = home_slot(key[q], cap);
                res[q] = kInvalidRow;
                if (j < it.s1) {
                    if (key[q] != kEmptyKey)
                        pend |= 1u << q;
                    else if (side_head != kInvalidRow)
                        res[q] = kSideSlot;
                }
            }
            while (pend) {
                uint64_t ck[LB];
#pragma unroll
                for (int q = 0; q < LB; ++q)
                    ck[q] = tkeys[pos[q]];
#pragma unroll
                for (int q = 0; q < LB; ++q) {
                    const bool p = (pend >> q) & 1u;
                    const bool hit = ck[q] == key[q];
                    const bool stop = hit || ck[q] == kEmptyKey;
                    res[q] = (p && hit) ? pos[q] : res[q];
                    pend &= ~((p && stop ? 1u : 0u) << q);
                    pos[q] = (p && !stop) ? (pos[q] + 1 == cap ? 0u : pos[q] + 1) : pos[q];
                }
            }
#pragma unroll
            for (int q = 0; q < LB; ++q) {
                const int k = g * LB + q;
                const uint32_t c = res[q] == kInvalidRow ? 0u : res[q] + 1;
                if (k % 2 == 0)
                    code[k / 2] = c;
                else
                    code[k / 2] |= c << 16;
            }
            if (g + 1 < K / LB) {
#pragma unroll
                for (int q = 0; q < LB; ++q)
                    key[q] = nkey[q];
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint64_t m = __ballot(((code[k / 2] >> (16 * (k % 2))) & 0xFFFFu) != 0);
            if (lane == 0)
                slot_off[k][wave] = static_cast<uint32_t>(__popcll(m));
        }
        __syncthreads();
        if (wave == 0) {
            // lane k: exclusive offsets of slot k's wave runs, slots in order
            uint32_t pre[W];
            uint32_t tot = 0;
#pragma unroll
            for (int ww = 0; ww < W; ++ww) {
                pre[ww] = tot;
                tot += lane < static_cast<uint32_t>(K) ? slot_off[lane < K ? lane : 0][ww] : 0u;
            }
            const uint32_t incl = wave_incl_scan(tot);
            if (lane < static_cast<uint32_t>(K)) {
#pragma unroll
                for (int ww = 0; ww < W; ++ww)
                    slot_off[lane][ww] = incl - tot + pre[ww];
            }
            if (lane == 63) {
                sr_total = incl;
                if (incl)
                    sr_base = atomicAdd(cursor, static_cast<unsigned long long>(incl));
            }
        }
        __syncthreads();
        if (sr_total != 0) {
            const uint64_t rb = sr_base;
            if constexpr (NARROW) {
#pragma unroll
                for (int r = 0; r < K / U; ++r) {
                    if (sb + r * U * THREADS >= it.s1)
                        continue; // uniform: rounds past the chunk's end
                    uint32_t ridx[U], sidx[U], opos[U];
                    uint64_t okey[U];
                    bool on[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = r * U + u;
                        const uint32_t c = (code[k / 2] >> (16 * (k % 2))) & 0xFFFFu;
                        const uint64_t m = __ballot(c != 0);
                        opos[u] = slot_off[k][wave] + mbcnt(m);
                        on[u] = c != 0 && rb + opos[u] < out_cap;
                        sidx[u] = k * THREADS + threadIdx.x;
                        uint32_t row = 0;
                        okey[u] = 0;
                        if (on[u]) {
                            if (c - 1 == kSideSlot) {
                                okey[u] = kEmptyKey;
                                row = side_head;
                            } else {
                                okey[u] = tkeys[c - 1];
                                row = trows[c - 1];
                            }
                        }
                        ridx[u] = row;
                    }
                    uint64_t *okeys = a.out.keys + rb;
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (on[u])
                            st_at<uint64_t>(okeys, opos[u] * 8u, okey[u]);
                    move_set<U>(a.r_cols, it.r0, a.out.left, rb, ridx, opos, on);
                    uint32_t srow[U];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        srow[u] = a.s_rowid[sb + (on[u] ? sidx[u] : 0u)];
                    gather_set_abs<U>(a.s_src, a.out.right, rb, srow, opos, on);
                }
            } else {
                // slot k: output row and matched table slot; the output key
                // is the slot's key (from LDS: no memory read), then the slot
                // becomes the build row for the column sweeps
                uint32_t meta[K];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const uint32_t c = (code[k / 2] >> (16 * (k % 2))) & 0xFFFFu;
                    const uint64_t m = __ballot(c != 0);
                    const uint32_t opos = slot_off[k][wave] + mbcnt(m);
                    const bool on = c != 0 && rb + opos < out_cap;
                    meta[k] = on ? (0x80000000u | (opos << 16) | (c - 1)) : 0u;
                }
                uint64_t *okeys = a.out.keys + rb;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (meta_on(meta[k])) {
                        const uint32_t slot = meta_row(meta[k]);
                        st_at<uint64_t>(okeys, meta_pos(meta[k]) * 8u, slot == kSideSlot ? kEmptyKey : tkeys[slot]);
                    }
                }
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const uint32_t slot = meta_row(meta[k]);
                    const uint32_t row = slot == kSideSlot ? side_head : trows[meta_on(meta[k]) ? slot : 0u];
                    meta[k] = (meta[k] & 0xFFFF0000u) | row;
                }
                const uint32_t none[K] = {};
                sweep_set<THREADS, K, kRowMeta>(a.r_cols, it.r0, a.out.left, rb, meta, none);
                sweep_set<THREADS, K, kRowProbe>(a.s_cols, sb, a.out.right, rb, meta, none);
            }
        }
        __syncthreads(); // slot_off and sr_* are rewritten by the next super-round
    }
}

// One workgroup per leaf: builds the leaf's fused-join table in LDS and
// writes it to the leaf's slot of the image as 16-byte (key, RowRefCell)
// cells {key, leaf, leaf-local row} (src/HashTable.h:35-40).
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_leaf_table_image(LeafJoinArgs a, unsigned char *__restrict__ img,
                                                              uint64_t stride, uint32_t *__restrict__ side,
                                                              uint32_t *__restrict__ dup_any) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t side_head;
    __shared__ uint32_t dup;
    Item it{};
    it.leaf = blockIdx.x;
    it.r0 = a.r_off[it.leaf];
    it.nr = a.r_off[it.leaf + 1] - it.r0;
    if (it.nr == 0 || it.nr > a.lds_max_rows) { // uniform
        if (threadIdx.x == 0)
            side[it.leaf] = kInvalidRow; // no table: probe rows find nothing
        return;
    }
    const uint32_t cap = table_cap(it.nr, a.load_pct);
    build_fused_table<THREADS>(a, it, cap, reinterpret_cast<uint64_t *>(smem),
                               reinterpret_cast<uint16_t *>(smem + static_cast<size_t>(cap) * 8), side_head, dup);
    const uint64_t *tkeys = reinterpret_cast<const uint64_t *>(smem);
    const uint16_t *trows = reinterpret_cast<const uint16_t *>(smem + static_cast<size_t>(cap) * 8);
    Cell16 *dst = reinterpret_cast<Cell16 *>(img + static_cast<uint64_t>(it.leaf) * stride);
    for (uint32_t i = threadIdx.x; i < cap; i += THREADS) {
        Cell16 c;
        c.key = tkeys[i];
        c.block_no = it.leaf;
        c.row_no = c.key == kEmptyKey ? kInvalidRow : trows[i];
        dst[i] = c;
    }
    if (threadIdx.x == 0) {
        side[it.leaf] = side_head;
        if (dup)
            atomicOr(dup_any, 1u);
    }
}

// Join of partitioned probe rows against prebuilt leaf tables (16-byte cells
// in HBM, no repeated build key): a work item is a fixed range of THREADS*K
// probe rows, whatever leaves they fall in. Each row's leaf is the top
// leaf_bits of its intHash64, the bits the radix passes consumed, so no leaf
// lookup table is needed; the rows of a leaf are adjacent, so neighbouring
// lanes read the same table. Lookups advance LB keys in lock step; the work
// item's output rows are reserved with one atomic add and laid out slot-major
// as in the fused join; then every column is written in pipelined sweeps over
// the lane's slots, as in the fused join.
template <int THREADS, int K>
__global__ __launch_bounds__(THREADS) void k_prebuilt_join(LeafJoinArgs a, uint64_t n_rows, uint32_t leaf_bits) {
    static_assert(K <= 64 && THREADS * K <= 0x8000, "K slots per lane; output rows of an item fit the metadata");
    constexpr int W = THREADS / 64;
    constexpr int LB = 4;
    static_assert(K % LB == 0, "lookup batches");
    __shared__ uint32_t slot_off[K][W];
    __shared__ uint32_t sr_total;
    __shared__ unsigned long long sr_base;
    const uint64_t sb = static_cast<uint64_t>(blockIdx.x) * THREADS * K;
    if (sb >= n_rows)
        return; // uniform
    const uint64_t s1 = min(n_rows, sb + static_cast<uint64_t>(THREADS) * K);
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    uint32_t ref[K]; // matched build row + 1 (build-side row index), 0: none
#pragma unroll
    for (int g = 0; g < K / LB; ++g) {
        uint64_t key[LB], base[LB];
        uint32_t pos[LB], cap[LB], r0[LB], res[LB];
        uint32_t pend = 0;
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            const uint64_t j = sb + static_cast<uint64_t>((g * LB + q) * THREADS + threadIdx.x);
            key[q] = a.s_keys[min(j, s1 - 1)]; // rows past the range are masked below
        }
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            const uint64_t j = sb + static_cast<uint64_t>((g * LB + q) * THREADS + threadIdx.x);
            const uint32_t leaf = leaf_bits ? static_cast<uint32_t>(int_hash64(key[q]) >> (64 - leaf_bits)) : 0u;
            r0[q] = a.r_off[leaf];
            const uint32_t nr = a.r_off[leaf + 1] - r0[q];
            cap[q] = table_cap(nr, a.load_pct);
            base[q] = static_cast<uint64_t>(leaf) * a.img_stride;
            pos[q] = home_slot(key[q], cap[q]);
            res[q] = kInvalidRow;
            if (j < s1) {
                if (key[q] == kEmptyKey)
                    res[q] = a.img_side[leaf];
                else if (nr > 0)
                    pend |= 1u << q;
            }
        }
        while (pend) {
            Cell16 c[LB];
#pragma unroll
            for (int q = 0; q < LB; ++q)
                c[q] = *reinterpret_cast<const Cell16 *>(a.table_img + base[q] + static_cast<uint64_t>(pos[q]) * sizeof(Cell16));
#pragma unroll
            for (int q = 0; q < LB; ++q) {
                const bool p = (pend >> q) & 1u;
                const bool hit = c[q].key == key[q];
                const bool stop = hit || c[q].key == kEmptyKey;
                res[q] = (p && hit) ? c[q].row_no : res[q];
                pend &= ~((p && stop ? 1u : 0u) << q);
                pos[q] = (p && !stop) ? (pos[q] + 1 == cap[q] ? 0u : pos[q] + 1) : pos[q];
            }
        }
#pragma unroll
        for (int q = 0; q < LB; ++q)
            ref[g * LB + q] = res[q] == kInvalidRow ? 0u : r0[q] + res[q] + 1;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint64_t m = __ballot(ref[k] != 0);
        if (lane == 0)
            slot_off[k][wave] = static_cast<uint32_t>(__popcll(m));
    }
    __syncthreads();
    if (wave == 0) {
        uint32_t pre[W];
        uint32_t tot = 0;
#pragma unroll
        for (int ww = 0; ww < W; ++ww) {
            pre[ww] = tot;
            tot += lane < static_cast<uint32_t>(K) ? slot_off[lane < K ? lane : 0][ww] : 0u;
        }
        const uint32_t incl = wave_incl_scan(tot);
        if (lane < static_cast<uint32_t>(K)) {
#pragma unroll
            for (int ww = 0; ww < W; ++ww)
                slot_off[lane][ww] = incl - tot + pre[ww];
        }
        if (lane == 63) {
            sr_total = incl;
            if (incl)
                sr_base = atomicAdd(a.out.cursor, static_cast<unsigned long long>(incl));
        }
    }
    __syncthreads();
    if (sr_total == 0)
        return; // uniform
    const uint64_t rb = sr_base;
    // slot k: output row (relative to rb) and matched build row
    uint32_t meta[K], brow[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint64_t m = __ballot(ref[k] != 0);
        const uint32_t o = slot_off[k][wave] + mbcnt(m);
        const bool on = ref[k] != 0 && rb + o < a.out.cap;
        meta[k] = on ? (0x80000000u | (o << 16)) : 0u;
        brow[k] = on ? ref[k] - 1 : 0u;
    }
    // the output key is the probe key (equal to the matched build key)
    sweep_col<THREADS, K, kRowProbe, uint64_t>(a.s_keys + sb, a.out.keys + rb, meta, brow, threadIdx.x);
    sweep_set<THREADS, K, kRowArray>(a.r_cols, 0, a.out.left, rb, meta, brow);
    sweep_set<THREADS, K, kRowProbe>(a.s_cols, sb, a.out.right, rb, meta, brow);
}

// ---------------------------------------------------------------- HBM tables

// Build columns of one-row keys from their wide cells: column c of cell ci[u]
// (byte (ci << shift) + its inline offset) to output row pos[u].
template <int U, class T>
__device__ __forceinline__ void gather_cell_col(const unsigned char *src, uint32_t shift, T *__restrict__ dst,
                                                const uint64_t (&ci)[U], const uint64_t (&pos)[U], const bool (&on)[U]) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
        v[u] = *reinterpret_cast<const T *>(src + ((on[u] ? ci[u] : 0u) << shift));
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (on[u])
            dst[pos[u]] = v[u];
}

template <int U>
__device__ __forceinline__ void gather_cells(const GlobalHT &ht, const ColSet &in, const OutColSet &out,
                                             const uint64_t (&ci)[U], const uint64_t (&pos)[U], const bool (&on)[U]) {
    const unsigned char *cells = reinterpret_cast<const unsigned char *>(ht.cells);
    const uint32_t shift = ht.cell_shift;
    uint32_t off = 16;
    for (uint32_t c = 0; c < in.n; ++c) {
        const uint32_t t = col_type(in, c);
        const uint32_t w = type_size(t);
        off = ht_pay_at(off, w);
        const unsigned char *src = cells + off;
        switch (t) {
        case 0: gather_cell_col<U>(src, shift, optr<uint8_t>(out, c), ci, pos, on); break;
        case 1: gather_cell_col<U>(src, shift, optr<uint16_t>(out, c), ci, pos, on); break;
        case 2: gather_cell_col<U>(src, shift, optr<uint32_t>(out, c), ci, pos, on); break;
        case 3: gather_cell_col<U>(src, shift, optr<uint64_t>(out, c), ci, pos, on); break;
        default: gather_cell_col<U>(src, shift, optr<ulonglong2>(out, c), ci, pos, on); break;
        }
        off += w;
    }
}

// Element of type t at byte address p (two 64-bit halves for 16 bytes).
__device__ __forceinline__ void load_at(const unsigned char *p, uint32_t t, uint64_t &lo, uint64_t &hi) {
    hi = 0;
    switch (t) {
    case 0: lo = *p; break;
    case 1: lo = *reinterpret_cast<const uint16_t *>(p); break;
    case 2: lo = *reinterpret_cast<const uint32_t *>(p); break;
    case 3: lo = *reinterpret_cast<const uint64_t *>(p); break;
    default: {
        const ulonglong2 v = *reinterpret_cast<const ulonglong2 *>(p);
        lo = v.x;
        hi = v.y;
    }
    }
}

// Loads column j of a round's one-match rows (see emit_one_rows); pay_off is
// the inline offset of build column j in a wide cell, advanced past it.
template <int U>
__device__ __forceinline__ void load_one_col(uint32_t j, const unsigned char *cells, uint32_t shift, const ColSet &rc,
                                             const ColSet &sc, const uint64_t (&ci)[U], const bool (&fromcell)[U],
                                             const ulonglong2 (&pay)[U], const uint32_t (&ridx)[U],
                                             const uint32_t (&sidx)[U], const bool (&on)[U], uint32_t &pay_off,
                                             uint64_t (&lo)[U], uint64_t (&hi)[U]) {
    if (j < rc.n) {
        const uint32_t t = col_type(rc, j);
        const uint32_t w = type_size(t);
        pay_off = ht_pay_at(pay_off, w);
        const unsigned char *col = static_cast<const unsigned char *>(col_ptr(rc, j));
        if (shift == 5) {
            // 32-byte cells: the column from the payload already in registers
            // (bytes [pay_off - 16, + w) of the 16 inline bytes, aligned to w)
            const uint32_t o = pay_off - 16;
            const uint64_t m = w >= 8 ? ~0ull : (1ull << (8 * w)) - 1;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (fromcell[u]) {
                    const uint64_t half = o >= 8 ? pay[u].y : pay[u].x;
                    lo[u] = w == 16 ? pay[u].x : (half >> (8 * (o & 7))) & m;
                    hi[u] = w == 16 ? pay[u].y : 0;
                } else {
                    load_at(col + static_cast<uint64_t>(on[u] ? ridx[u] : 0u) * w, t, lo[u], hi[u]);
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const unsigned char *p = fromcell[u] ? cells + (ci[u] << shift) + pay_off
                                                     : col + static_cast<uint64_t>(on[u] ? ridx[u] : 0u) * w;
                load_at(p, t, lo[u], hi[u]);
            }
        }
        pay_off += w;
    } else {
        const uint32_t c = j - rc.n;
        const uint32_t t = col_type(sc, c);
        const void *col = col_ptr(sc, c);
#pragma unroll
        for (int u = 0; u < U; ++u)
            load_elem(col, t, on[u] ? sidx[u] : 0u, lo[u], hi[u]);
    }
}

// The build and probe columns of a round's one-match rows, column after
// column, each column's loads issued before the previous column's stores (a
// load issued after a store waits for it: one counter orders a wave's memory
// operations). Build column c of row u comes from wide cell ci[u] when
// fromcell[u] (from its payload pay[u], already loaded, for 32-byte cells),
// else from the build columns at ridx[u]; probe columns from sidx.
template <int U>
__device__ __forceinline__ void emit_one_rows(const GlobalHT &ht, const ColSet &rc, const ColSet &sc, const OutSink &out,
                                              const uint64_t (&ci)[U], const bool (&fromcell)[U],
                                              const ulonglong2 (&pay)[U], const uint32_t (&ridx)[U],
                                              const uint32_t (&sidx)[U], const uint64_t (&opos)[U], const bool (&on)[U]) {
    const uint32_t nc = rc.n + sc.n;
    if (nc == 0)
        return;
    const unsigned char *cells = reinterpret_cast<const unsigned char *>(ht.cells);
    const uint32_t shift = ht.cell_shift;
    uint32_t pay_off = 16;
    uint64_t lo[U], hi[U];
    load_one_col<U>(0, cells, shift, rc, sc, ci, fromcell, pay, ridx, sidx, on, pay_off, lo, hi);
    for (uint32_t j = 0; j < nc; ++j) {
        uint64_t nlo[U], nhi[U];
        if (j + 1 < nc)
            load_one_col<U>(j + 1, cells, shift, rc, sc, ci, fromcell, pay, ridx, sidx, on, pay_off, nlo, nhi);
        uint32_t t;
        void *dst;
        if (j < rc.n) { // no reference chosen between the two column sets: it would copy both to scratch
            t = col_type(rc, j);
            dst = col_ptr(out.left, j);
        } else {
            t = col_type(sc, j - rc.n);
            dst = col_ptr(out.right, j - rc.n);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (on[u])
                store_elem(dst, t, opos[u], lo[u], hi[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            lo[u] = nlo[u];
            hi[u] = nhi[u];
        }
    }
}

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

// Bit of an HBM table cell's row_no set once a second build row of the cell's
// key is inserted: probes of keys without it take the head as the only match
// and skip the chain walk (one random read of next[] less per matched row).
// Row numbers stay below 2^31 (block rows, leaf rows).
constexpr uint32_t kRefMore = 0x80000000u;

__device__ __forceinline__ uint32_t ref_row(const HtArgs &a, uint64_t ref) {
    const uint32_t bno = static_cast<uint32_t>(ref);
    const uint32_t rno = static_cast<uint32_t>(ref >> 32) & ~kRefMore;
    if (bno == kInvalidRow)
        return kInvalidRow;
    if (a.mode == kHtLeaves)
        return a.ht.build_seg_off[bno] + rno;
    return bno * a.ht.block_rows + rno;
}

__device__ __forceinline__ uint32_t cell_shift_of(const GlobalHT &ht) { return ht.cell_shift ? ht.cell_shift : 4u; }
__device__ __forceinline__ Cell16 *cell_at(const GlobalHT &ht, uint64_t i) {
    return reinterpret_cast<Cell16 *>(reinterpret_cast<unsigned char *>(ht.cells) + (i << cell_shift_of(ht)));
}

// Every 16-byte word of the table is written (a wide cell's inline columns
// as zeros): whole cache lines leave L2, not lines with holes in them.
__global__ void k_ht_init(GlobalHT ht, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    const uint32_t wsh = cell_shift_of(ht) - 4; // log2 of 16-byte words per cell
    const uint64_t words = n << wsh;
    Cell16 *w16 = ht.cells;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) {
        Cell16 c;
        const bool head = (i & ((1u << wsh) - 1)) == 0;
        c.key = head ? kEmptyKey : 0;
        c.block_no = head ? kInvalidRow : 0;
        c.row_no = 0;
        w16[i] = c;
    }
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g < ht.n_tables)
        ht.side_head[g] = kInvalidRow;
}

// Copies build row `row` of the payload columns into a wide cell.
__device__ __forceinline__ void put_payload(const GlobalHT &ht, Cell16 *cell, uint64_t row) {
    unsigned char *base = reinterpret_cast<unsigned char *>(cell);
    uint32_t off = 16;
#pragma unroll 1
    for (uint32_t c = 0; c < ht.pay.n; ++c) {
        const uint32_t t = col_type(ht.pay, c);
        const uint32_t w = type_size(t);
        off = ht_pay_at(off, w);
        const void *src = col_ptr(ht.pay, c);
        switch (t) {
        case 0: *reinterpret_cast<uint8_t *>(base + off) = static_cast<const uint8_t *>(src)[row]; break;
        case 1: *reinterpret_cast<uint16_t *>(base + off) = static_cast<const uint16_t *>(src)[row]; break;
        case 2: *reinterpret_cast<uint32_t *>(base + off) = static_cast<const uint32_t *>(src)[row]; break;
        case 3: *reinterpret_cast<uint64_t *>(base + off) = static_cast<const uint64_t *>(src)[row]; break;
        default:
            *reinterpret_cast<ulonglong2 *>(base + off) = static_cast<const ulonglong2 *>(src)[row];
            break;
        }
        off += w;
    }
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
            Cell16 *c = cell_at(a.ht, base + pos);
            const unsigned long long old = atomicCAS(reinterpret_cast<unsigned long long *>(&c->key),
                                                     static_cast<unsigned long long>(kEmptyKey),
                                                     static_cast<unsigned long long>(k));
            if (old == kEmptyKey || old == k) {
                prev = atomicExch(reinterpret_cast<unsigned long long *>(&c->block_no), static_cast<unsigned long long>(myref));
                // every exchange that displaces a row is followed by its own
                // OR, so the flag ends set exactly when the key has 2+ rows
                if (static_cast<uint32_t>(prev) != kInvalidRow) {
                    atomicOr(&c->row_no, kRefMore);
                    if (a.ht.dups)
                        *a.ht.dups = 1u;
                } else if (a.ht.pay.n)
                    put_payload(a.ht, c, row); // the key's first row: read only while it stays the only one
                break;
            }
            pos = (pos + 1) & mask;
        }
        a.ht.next[row] = ref_row(a, prev);
    }
}

// One thread per cell: the row count of a flagged cell's chain.
__global__ __launch_bounds__(256) void k_ht_group_count(HtArgs a, uint64_t total_cells, uint32_t *__restrict__ cnt) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i <= total_cells; i += stride) {
        uint32_t len = 0;
        if (i < total_cells) {
            const Cell16 c = *cell_at(a.ht, i);
            if (c.key != kEmptyKey && (c.row_no & kRefMore))
                for (uint32_t r = ref_row(a, static_cast<uint64_t>(c.block_no) | (static_cast<uint64_t>(c.row_no) << 32));
                     r != kInvalidRow; r = a.ht.next[r])
                    ++len;
        }
        cnt[i] = len;
    }
}

// One thread per cell: a flagged cell's chain rows (newest first, the order
// a chain walk emits them) listed at their group's place in rows[], then the
// cell rewritten as {group start, row count | kRefMore}. The build columns
// follow in one gather over rows[].
__global__ __launch_bounds__(256) void k_ht_group_fill(HtArgs a, uint64_t total_cells, const uint32_t *__restrict__ goff,
                                                       uint32_t *__restrict__ rows) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total_cells; i += stride) {
        Cell16 *cp = cell_at(a.ht, i);
        const Cell16 c = *cp;
        if (c.key == kEmptyKey || !(c.row_no & kRefMore))
            continue;
        const uint32_t g0 = goff[i];
        uint32_t k = g0;
        for (uint32_t r = ref_row(a, static_cast<uint64_t>(c.block_no) | (static_cast<uint64_t>(c.row_no) << 32));
             r != kInvalidRow; r = a.ht.next[r], ++k)
            rows[k] = r;
        cp->block_no = g0;
        cp->row_no = (k - g0) | kRefMore;
    }
}

__device__ __forceinline__ uint32_t ht_lookup(const HtArgs &a, uint64_t k, uint32_t table) {
    const uint64_t h = int_hash64(k);
    uint64_t base, mask, pos;
    ht_home(a, h, table, base, mask, pos);
    if (k == kEmptyKey)
        return a.ht.side_head[table];
    while (true) {
        const Cell16 c = *cell_at(a.ht, base + pos);
        if (c.key == k)
            return ref_row(a, static_cast<uint64_t>(c.block_no) | (static_cast<uint64_t>(c.row_no) << 32));
        if (c.key == kEmptyKey)
            return kInvalidRow;
        pos = (pos + 1) & mask;
    }
}

// LDS for the probe columns of a round's grouped slots.
constexpr size_t kHtGroupProbeLds = 8 * 1024;

// Probe + emit against HBM tables (CHJ, oversize leaves). Per round of
// THREADS*U probe rows: all U lookups of a thread advance in lock step (one
// probe step each per iteration, their cell loads in flight together), chain
// lengths likewise, then one output reservation per workgroup. Rows with one
// match are written column by column with every load of a column issued
// before its stores; rows with several matches (repeated build keys) walk
// their chain row by row. Output row = probe key | build columns | probe
// columns (src/JoinOps.h:77-125).
//
// With grouped chains (CHJ) a repeated key's cell names its group, so its rows
// need no walk: the workgroup writes the round's group rows together, each
// thread taking consecutive output rows and finding their slot by binary
// search over the slots' group-row ends (as the two-phase leaf emit does).
template <int THREADS, int U>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(6))) void k_ht_probe_emit(HtArgs a, ColSet rc, ColSet sc, OutSink out) {
    constexpr int kSlots = THREADS * U;
    __shared__ uint32_t tmp[2 * (THREADS / 64 + 1)];
    __shared__ unsigned long long out_base;
    // the round's grouped slots, listed compactly (see the group rows below)
    __shared__ uint32_t g_end[kSlots], g_st[kSlots], g_row[kSlots];
    __shared__ uint64_t g_key[kSlots];
    __shared__ __align__(16) unsigned char g_probe[kHtGroupProbeLds];
    const uint64_t n = ht_total_rows(a);
    const uint64_t per_round = static_cast<uint64_t>(gridDim.x) * THREADS * U;
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * THREADS * U; base < n; base += per_round) {
        uint64_t key[U], prow[U], cb[U], cm[U], cp[U];
        uint32_t head[U], len[U], gst[U];
        uint32_t pend = 0, many = 0; // many: rows whose chain must be walked
        uint32_t grpm = 0;           // rows whose key's rows are one group (grouped chains)
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t g = base + static_cast<uint64_t>(u) * THREADS + threadIdx.x;
            uint64_t row;
            uint32_t table, leaf;
            ht_locate(a, min(g, n - 1), row, table, leaf); // rows past the end are masked below
            prow[u] = row;
            key[u] = a.rows.keys[row];
            ht_home(a, int_hash64(key[u]), table, cb[u], cm[u], cp[u]);
            head[u] = kInvalidRow;
            if (g < n) {
                if (key[u] == kEmptyKey) {
                    head[u] = a.ht.side_head[table];
                    many |= 1u << u;
                } else {
                    pend |= 1u << u;
                }
            }
        }
        // Two cells per step when the next cell shares the probe position's
        // 128-byte line (loaded together, one line fetched): a wave waits for
        // its longest probe sequence, and at load factor 0.5 most end within
        // the first line.
        const uint32_t sh = cell_shift_of(a.ht);
        while (pend) {
            Cell16 c[U], c2[U];
            bool two[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t i0 = cb[u] + cp[u];
                const uint64_t i1 = cb[u] + ((cp[u] + 1) & cm[u]);
                two[u] = ((i0 << sh) >> 7) == ((i1 << sh) >> 7);
                c[u] = *cell_at(a.ht, i0);
                c2[u] = *cell_at(a.ht, two[u] ? i1 : i0); // no second line fetched
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!((pend >> u) & 1u))
                    continue;
                if (two[u] && c[u].key != key[u] && c[u].key != kEmptyKey) {
                    // the first cell neither holds the key nor ends the sequence
                    c[u] = c2[u];
                    cp[u] = (cp[u] + 1) & cm[u];
                }
                if (c[u].key == key[u]) {
                    if ((c[u].row_no & kRefMore) && a.ht.grouped) {
                        grpm |= 1u << u;
                        gst[u] = c[u].block_no;              // group start
                        len[u] = c[u].row_no & ~kRefMore;    // group rows
                    } else {
                        head[u] = ref_row(a, static_cast<uint64_t>(c[u].block_no) | (static_cast<uint64_t>(c[u].row_no) << 32));
                        if (c[u].row_no & kRefMore)
                            many |= 1u << u;
                    }
                    pend &= ~(1u << u); // cb + cp now names the matched cell
                } else if (c[u].key == kEmptyKey) {
                    pend &= ~(1u << u);
                } else {
                    cp[u] = (cp[u] + 1) & cm[u];
                }
            }
        }
        // 32-byte cells: the matched cell's inline build columns, read while
        // its line is still cached (by the time the round's rows are written
        // the line has left L2)
        ulonglong2 pay[U];
        if (sh == 5) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                pay[u] = *reinterpret_cast<const ulonglong2 *>(reinterpret_cast<const unsigned char *>(cell_at(a.ht, cb[u] + cp[u])) + 16);
        }
        // chain lengths, all chains of the thread one step per iteration;
        // a key with one build row has length 1 without a walk
        uint32_t x[U];
        uint32_t walking = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            x[u] = head[u];
            if ((grpm >> u) & 1u)
                continue; // len is the group's row count
            len[u] = 0;
            if (x[u] != kInvalidRow) {
                if ((many >> u) & 1u)
                    walking |= 1u << u;
                else
                    len[u] = 1;
            }
        }
        while (walking) {
            uint32_t nx[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                nx[u] = a.ht.next[(walking >> u) & 1u ? x[u] : 0u];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!((walking >> u) & 1u))
                    continue;
                ++len[u];
                x[u] = nx[u];
                if (x[u] == kInvalidRow)
                    walking &= ~(1u << u);
            }
        }
        // The round's output: its one-match and walked rows thread by thread,
        // then the rows of its grouped keys as one contiguous range.
        uint32_t cnt = 0, gcnt = 0, ne = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if ((grpm >> u) & 1u) {
                gcnt += len[u];
                ne += len[u] > 0 ? 1u : 0u;
            } else {
                cnt += len[u];
            }
        }
        uint32_t off, g_off, t_ow, g_total;
        block_excl_scan2<THREADS>(cnt, gcnt, tmp, off, g_off, t_ow, g_total);
        const uint32_t round_total = t_ow + g_total;
        if (round_total == 0)
            continue;
        if (threadIdx.x == 0)
            out_base = atomicAdd(out.cursor, static_cast<unsigned long long>(round_total));
        __syncthreads();
        // Past the sink's capacity (a count-only sink has none) the round's
        // rows are only counted. Uniform: every thread read the same out_base,
        // and thread 0 rewrites it only after the next round's scan barriers.
        if (out_base >= out.cap)
            continue;
        uint64_t opos[U];
        uint32_t ridx[U], sidx[U];
        bool one[U];
        uint64_t pos = out_base + off;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            opos[u] = pos;
            pos += (grpm >> u) & 1u ? 0u : len[u];
            one[u] = len[u] == 1 && !((grpm >> u) & 1u) && opos[u] < out.cap;
            ridx[u] = one[u] ? head[u] : 0u;
            sidx[u] = static_cast<uint32_t>(prow[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (one[u])
                out.keys[opos[u]] = key[u];
        {
            // one-row keys take their build columns from the matched wide
            // cell (just read); others (narrow cells, a one-row side chain of
            // key kEmptyKey) from the build columns
            uint64_t ci[U];
            bool fromcell[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                fromcell[u] = a.ht.pay.n && one[u] && !((many >> u) & 1u);
                ci[u] = fromcell[u] ? cb[u] + cp[u] : 0u;
            }
            emit_one_rows<U>(a.ht, rc, sc, out, ci, fromcell, pay, ridx, sidx, opos, one);
        }
        if (g_total > 0) { // uniform
            // Grouped slots with rows, in output order: slot k's rows end at
            // g_end[k] (relative to the group range), its key's build rows
            // start at g_st[k] of the grouped columns; its probe row and key,
            // and its probe columns when they fit, are kept in LDS.
            uint32_t n_seg;
            uint32_t k = block_excl_scan<THREADS>(ne, tmp, &n_seg);
            uint32_t nlds = 0; // probe columns staged: all or none
            {
                size_t used = 0;
                for (uint32_t c = 0; c < sc.n; ++c)
                    used += emit_rcol_bytes(kSlots, type_size(col_type(sc, c)));
                nlds = used <= kHtGroupProbeLds ? sc.n : 0u;
            }
            uint32_t e = g_off;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (((grpm >> u) & 1u) && len[u] > 0) {
                    e += len[u];
                    g_end[k] = e;
                    g_st[k] = gst[u];
                    g_row[k] = static_cast<uint32_t>(prow[u]);
                    g_key[k] = key[u];
                    size_t co = 0;
                    for (uint32_t c = 0; c < nlds; ++c) {
                        const uint32_t t = col_type(sc, c);
                        uint64_t lo, hi;
                        load_elem(col_ptr(sc, c), t, prow[u], lo, hi);
                        store_elem(g_probe + co, t, k, lo, hi);
                        co += emit_rcol_bytes(kSlots, type_size(t));
                    }
                    ++k;
                }
            }
            __syncthreads();
            // Output rows in aligned groups of four consecutive rows per
            // thread: a group's rows mostly belong to one slot's consecutive
            // build rows (one search per group, then a step to the next slot),
            // and each column of a full group leaves as one packed store.
            const uint64_t gbase = out_base + t_ow;
            const bool pack_ok = (reinterpret_cast<uint64_t>(out.keys) & 15) == 0 && cols_pack_aligned(out.left) &&
                                 cols_pack_aligned(out.right);
            const uint32_t lead = static_cast<uint32_t>(gbase & 3u);
            const uint64_t a0 = gbase - lead;
            const uint32_t ngroups = (lead + g_total + 3) / 4;
            for (uint32_t q0 = 0; q0 < ngroups; q0 += THREADS) {
                const uint32_t gi = q0 + threadIdx.x;
                const int64_t o0 = static_cast<int64_t>(gi) * 4 - lead; // group-range row of the first row
                uint32_t oc = static_cast<uint32_t>(min<int64_t>(max<int64_t>(o0, 0), g_total - 1));
                uint32_t sg = 0; // first slot ending after oc (the same number of steps in every lane)
                for (uint32_t m = n_seg; m > 0;) {
                    const uint32_t half = m >> 1;
                    if (g_end[sg + half] <= oc) {
                        sg += half + 1;
                        m -= half + 1;
                    } else {
                        m = half;
                    }
                }
                uint32_t bidx[4], slot[4], sidx[4];
                bool on[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t o = o0 + r;
                    on[r] = gi < ngroups && o >= 0 && o < static_cast<int64_t>(g_total) && gbase + o < out.cap;
                    if (on[r])
                        oc = static_cast<uint32_t>(o); // rows without output repeat the last row with one
                    sg += oc >= g_end[sg] ? 1u : 0u; // the row starts the next slot's rows
                    bidx[r] = g_st[sg] + (oc - (sg ? g_end[sg - 1] : 0u));
                    slot[r] = sg;
                    sidx[r] = nlds ? sg : g_row[sg]; // staged probe columns by slot, else by probe row
                }
                const uint64_t p = a0 + static_cast<uint64_t>(gi) * 4;
                const bool packed = pack_ok && on[0] && on[1] && on[2] && on[3];
                move4<uint64_t>(g_key, slot, out.keys, p, on, packed);
                move4_set(nullptr, 0, 0u, a.ht.grp, 0, out.left, bidx, p, on, packed);
                move4_set(g_probe, kSlots, nlds, sc, 0, out.right, sidx, p, on, packed);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (len[u] < 2 || ((grpm >> u) & 1u))
                continue;
            uint64_t p = opos[u];
            for (uint32_t r = head[u]; r != kInvalidRow; r = a.ht.next[r])
                emit_row(out, p++, key[u], rc, r, sc, prow[u]);
        }
        __syncthreads(); // out_base is rewritten by the next round
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
    // 16-byte cells, then next, gstart and glen_h (u16 per build row each)
    return static_cast<size_t>(table_cap(lds_max_rows, load_pct)) * sizeof(LdsCell) +
           3 * static_cast<size_t>(round_up16(lds_max_rows)) * 2;
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
    k_leaf_probe<T, U><<<checked_grid(grid8(max_work), T, "leaf_probe"), T, lds, s>>>(a);
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

uint32_t leaf_join_fused_chunk() { return kFusedChunkRows; }

void leaf_chain_slots(hipStream_t s, const LeafJoinArgs &a, uint32_t *chain_first, uint32_t *scan_tmp) {
    k_leaf_chain_slots<<<(a.n_leaves + 1 + 255) / 256, 256, 0, s>>>(a.r_off, a.work_first, a.n_leaves, chain_first);
    scan_u32(s, chain_first, a.n_leaves + 1, nullptr, 1, scan_tmp, nullptr, "join_plan");
}

void prebuilt_join(hipStream_t s, const LeafJoinArgs &a, uint64_t probe_rows, uint32_t leaf_bits, KernelTimer *tm) {
    constexpr int T = 256, K = 8;
    if (probe_rows == 0)
        return;
    TimedScope ts(tm, "leaf_join", s);
    k_prebuilt_join<T, K><<<checked_grid((probe_rows + T * K - 1) / (T * K), T, "prebuilt_join"), T, 0, s>>>(a, probe_rows,
                                                                                                          leaf_bits);
}

size_t leaf_table_bytes(uint32_t rows, uint32_t load_pct) { return static_cast<size_t>(table_cap(rows, load_pct)) * sizeof(Cell16); }

void leaf_table_images(hipStream_t s, const LeafJoinArgs &a, unsigned char *img, uint64_t stride, uint32_t *side,
                       uint32_t *dup, KernelTimer *tm) {
    if (a.n_leaves == 0)
        return;
    const size_t lds = fused_table_bytes(a.lds_max_rows, a.load_pct);
    TimedScope ts(tm, "leaf_table_image", s);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_leaf_table_image<512>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    k_leaf_table_image<512><<<a.n_leaves, 512, lds, s>>>(a, img, stride, side, dup);
}

void leaf_join_fused(hipStream_t s, const LeafJoinArgs &a, uint64_t max_work, KernelTimer *tm) {
    if (max_work == 0)
        return;
    const size_t lds = fused_table_bytes(a.lds_max_rows, a.load_pct);
    if (a.s_rowid) {
        TimedScope ts(tm, "leaf_join_narrow", s);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_leaf_join<512, kFusedRowsPerThread, 2, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        k_leaf_join<512, kFusedRowsPerThread, 2, true><<<checked_grid(grid8(max_work), 512, "leaf_join"), 512, lds, s>>>(a);
        return;
    }
    TimedScope ts(tm, "leaf_join", s);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_leaf_join<512, kFusedRowsPerThread, 2, false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    k_leaf_join<512, kFusedRowsPerThread, 2, false><<<checked_grid(grid8(max_work), 512, "leaf_join"), 512, lds, s>>>(a);
}

void leaf_emit(hipStream_t s, const LeafJoinArgs &a, uint64_t max_work, bool chains, KernelTimer *tm) {
    if (max_work == 0)
        return;
    const size_t lds_chains = emit_chain_lds(a.lds_max_rows);
    {
        TimedScope ts(tm, "leaf_emit", s);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_leaf_emit<kEmitThreads, kEmitSlots, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_chains));
        k_leaf_emit<kEmitThreads, kEmitSlots, false>
            <<<checked_grid(grid8(max_work), kEmitThreads, "leaf_emit"), kEmitThreads, lds_chains, s>>>(a);
    }
    if (chains) {
        TimedScope ts(tm, "leaf_emit_chains", s);
        // Build columns go to LDS when they fit beside the groups: the chain
        // rows then gather them from LDS instead of from random HBM lines.
        const size_t rbytes = emit_rcols_bytes(a.r_cols, a.lds_max_rows);
        LeafJoinArgs b = a;
        b.r_lds = a.r_cols.n > 0 && rbytes <= kEmitRcolLds ? 1u : 0u;
        // Each round's probe keys and the leading probe columns that fit
        // kEmitScolLds go to LDS.
        constexpr uint32_t kRoundRows = kEmitThreads * kEmitSlots;
        size_t sbytes = emit_rcol_bytes(kRoundRows, 8);
        b.s_lds = sbytes <= kEmitScolLds ? 1u : 0u;
        if (!b.s_lds)
            sbytes = 0;
        while (b.s_lds && b.s_lds - 1 < a.s_cols.n &&
               sbytes + emit_rcol_bytes(kRoundRows, type_size(a.s_cols.types[b.s_lds - 1])) <= kEmitScolLds)
            sbytes += emit_rcol_bytes(kRoundRows, type_size(a.s_cols.types[b.s_lds++ - 1]));
        constexpr size_t kLdsMax = 160 * 1024;
        const size_t fixed = lds_chains + kRoundRows * (4 + 2 + 2);
        if (fixed + (b.r_lds ? rbytes : 0) + sbytes > kLdsMax) {
            b.s_lds = 0; // probe rows gathered from memory
            sbytes = 0;
        }
        if (fixed + (b.r_lds ? rbytes : 0) > kLdsMax)
            b.r_lds = 0;
        const size_t lds = fixed + (b.r_lds ? rbytes : 0) + sbytes;
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_leaf_emit<kEmitThreads, kEmitSlots, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        k_leaf_emit<kEmitThreads, kEmitSlots, true>
            <<<checked_grid(grid8(max_work), kEmitThreads, "leaf_emit"), kEmitThreads, lds, s>>>(b);
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
    const uint64_t n = std::max<uint64_t>(total_cells << ((ht.cell_shift ? ht.cell_shift : 4u) - 4), ht.n_tables);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((n + 255) / 256, 256 * 64));
    k_ht_init<<<grid, 256, 0, s>>>(ht, total_cells);
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
    constexpr int T = 256, U = 2; // measured on 2^30 CHJ probes: faster than U = 4 or 8
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((max_rows + T * U - 1) / (T * U), 256 * 64));
    k_ht_probe_emit<T, U><<<grid, T, 0, s>>>(make_args(ht, mode, probe), r_cols, s_cols, out);
}

void ht_group_count(hipStream_t s, const GlobalHT &ht, HtMode mode, uint64_t total_cells, uint32_t *cnt, KernelTimer *tm) {
    TimedScope ts(tm, "ht_group", s);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((total_cells + 256) / 256, 256 * 64));
    HtRows r{};
    k_ht_group_count<<<grid, 256, 0, s>>>(make_args(ht, mode, r), total_cells, cnt);
}

void ht_group_fill(hipStream_t s, const GlobalHT &ht, HtMode mode, uint64_t total_cells, const uint32_t *goff,
                   uint32_t *rows, KernelTimer *tm) {
    if (total_cells == 0)
        return;
    TimedScope ts(tm, "ht_group", s);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((total_cells + 255) / 256, 256 * 64));
    HtRows r{};
    k_ht_group_fill<<<grid, 256, 0, s>>>(make_args(ht, mode, r), total_cells, goff, rows);
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

// ------------------------------------------------------------ group images

__host__ __device__ __forceinline__ size_t group_lds(uint32_t max_rows, uint32_t load_pct) {
    const size_t cap = table_cap(max_rows, load_pct);
    return cap * 8 + ((cap * 4 + 15) & ~size_t{15});
}

// One workgroup per leaf: counts the rows of each key in an LDS table, turns
// the counts into group starts (exclusive scan in slot order), writes the
// leaf's cells {key, start, count} and places every row in its key's group
// (perm[start + rank] = row). Empty-key rows form the leaf's last group.
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_group_image(const uint64_t *__restrict__ r_keys,
                                                         const uint32_t *__restrict__ r_off, uint32_t max_rows,
                                                         uint32_t load_pct, Cell16 *__restrict__ cells,
                                                         uint64_t stride, uint32_t *__restrict__ side,
                                                         uint32_t *__restrict__ perm) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t side_cnt, side_cur;
    __shared__ uint32_t scan_tmp[THREADS / 64 + 1];
    const uint32_t leaf = blockIdx.x;
    const uint32_t r0 = r_off[leaf], nr = r_off[leaf + 1] - r0;
    if (nr == 0 || nr > max_rows) { // uniform; the caller sizes max_rows to the largest leaf
        if (threadIdx.x == 0) {
            side[2 * leaf] = r0;
            side[2 * leaf + 1] = 0;
        }
        return;
    }
    const uint32_t cap = table_cap(nr, load_pct);
    uint64_t *tkeys = reinterpret_cast<uint64_t *>(smem);
    uint32_t *tcnt = reinterpret_cast<uint32_t *>(smem + static_cast<size_t>(cap) * 8);
    for (uint32_t i = threadIdx.x; i < cap; i += THREADS) {
        tkeys[i] = kEmptyKey;
        tcnt[i] = 0;
    }
    if (threadIdx.x == 0)
        side_cnt = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nr; i += THREADS) {
        const uint64_t k = r_keys[r0 + i];
        if (k == kEmptyKey) {
            atomicAdd(&side_cnt, 1u);
            continue;
        }
        uint32_t pos = home_slot(k, cap);
        while (true) {
            const unsigned long long old =
                atomicCAS(reinterpret_cast<unsigned long long *>(&tkeys[pos]),
                          static_cast<unsigned long long>(kEmptyKey), static_cast<unsigned long long>(k));
            if (old == kEmptyKey || old == k)
                break;
            pos = pos + 1 == cap ? 0u : pos + 1;
        }
        atomicAdd(&tcnt[pos], 1u);
    }
    __syncthreads();
    // slots in THREADS chunks of consecutive slots
    const uint32_t per = (cap + THREADS - 1) / THREADS;
    const uint32_t c0 = min(cap, threadIdx.x * per), c1 = min(cap, c0 + per);
    uint32_t sum = 0;
    for (uint32_t i = c0; i < c1; ++i)
        sum += tcnt[i];
    uint32_t total;
    uint32_t start = block_excl_scan<THREADS>(sum, scan_tmp, &total);
    Cell16 *dst = cells + static_cast<uint64_t>(leaf) * stride;
    for (uint32_t i = c0; i < c1; ++i) {
        const uint32_t n = tcnt[i];
        Cell16 c;
        c.key = tkeys[i];
        c.block_no = r0 + start;
        c.row_no = n;
        dst[i] = c;
        tcnt[i] = start; // the slot's placement cursor from here on
        start += n;
    }
    if (threadIdx.x == 0) {
        side[2 * leaf] = r0 + total;
        side[2 * leaf + 1] = side_cnt;
        side_cur = total;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nr; i += THREADS) {
        const uint64_t k = r_keys[r0 + i];
        uint32_t at;
        if (k == kEmptyKey) {
            at = atomicAdd(&side_cur, 1u);
        } else {
            uint32_t pos = home_slot(k, cap);
            while (tkeys[pos] != k)
                pos = pos + 1 == cap ? 0u : pos + 1;
            at = atomicAdd(&tcnt[pos], 1u);
        }
        perm[r0 + at] = r0 + i;
    }
}

__global__ __launch_bounds__(256) void k_gather_perm(ColSet in, OutColSet out, const uint32_t *__restrict__ perm,
                                                     uint64_t rows) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < rows; i += stride) {
        const uint32_t src = perm[i];
        for (uint32_t c = 0; c < in.n; ++c) {
            const uint32_t t = col_type(in, c);
            uint64_t lo, hi;
            load_elem(col_ptr(in, c), t, src, lo, hi);
            store_elem(col_ptr(out, c), t, i, lo, hi);
        }
    }
}

// (start, count) of LB keys in the group images, looked up in lock step; a
// key outside `valid`, a miss, or a leaf without rows gives count 0.
template <int LB>
__device__ __forceinline__ void group_find(const GroupImage &g, const uint64_t (&key)[LB], const bool (&valid)[LB],
                                           uint32_t (&start)[LB], uint32_t (&count)[LB]) {
    uint32_t pos[LB], cap[LB];
    uint64_t base[LB];
    uint32_t pend = 0;
#pragma unroll
    for (int q = 0; q < LB; ++q) {
        const uint32_t leaf = g.leaf_bits ? static_cast<uint32_t>(int_hash64(key[q]) >> (64 - g.leaf_bits)) : 0u;
        const uint32_t r0 = g.r_off[leaf];
        const uint32_t nr = g.r_off[leaf + 1] - r0;
        cap[q] = table_cap(nr, g.load_pct);
        base[q] = static_cast<uint64_t>(leaf) * g.stride;
        pos[q] = home_slot(key[q], cap[q]);
        start[q] = 0;
        count[q] = 0;
        if (valid[q] && nr > 0) {
            if (key[q] == kEmptyKey) {
                start[q] = g.side[2 * leaf];
                count[q] = g.side[2 * leaf + 1];
            } else {
                pend |= 1u << q;
            }
        }
    }
    while (pend) {
        Cell16 c[LB];
#pragma unroll
        for (int q = 0; q < LB; ++q)
            c[q] = g.cells[base[q] + pos[q]];
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            const bool p = (pend >> q) & 1u;
            const bool hit = c[q].key == key[q];
            const bool stop = hit || c[q].key == kEmptyKey;
            start[q] = (p && hit) ? c[q].block_no : start[q];
            count[q] = (p && hit) ? c[q].row_no : count[q];
            pend &= ~((p && stop ? 1u : 0u) << q);
            pos[q] = (p && !stop) ? (pos[q] + 1 == cap[q] ? 0u : pos[q] + 1) : pos[q];
        }
    }
}

template <int THREADS, int K>
__global__ __launch_bounds__(THREADS) void k_group_count(GroupImage g, const uint64_t *__restrict__ keys, uint64_t rows,
                                                         unsigned long long *__restrict__ total) {
    constexpr int LB = 4;
    static_assert(K % LB == 0, "lookup batches");
    const uint64_t sb = static_cast<uint64_t>(blockIdx.x) * THREADS * K;
    uint64_t sum = 0;
#pragma unroll
    for (int gq = 0; gq < K / LB; ++gq) {
        uint64_t key[LB];
        bool valid[LB];
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            const uint64_t j = sb + static_cast<uint64_t>((gq * LB + q) * THREADS + threadIdx.x);
            valid[q] = j < rows;
            key[q] = keys[min(j, rows - 1)];
        }
        uint32_t st[LB], cn[LB];
        group_find<LB>(g, key, valid, st, cn);
#pragma unroll
        for (int q = 0; q < LB; ++q)
            sum += cn[q];
    }
    sum = wave_sum_u64(sum);
    if (lane_id() == 0 && sum != 0)
        atomicAdd(total, static_cast<unsigned long long>(sum));
}

// Rows o of a work item's output, G per lane at a time: entry e (probe row
// sb + e) and build row b of each; every column moves G rows with the loads
// issued before the stores.
// LDS for a group-join work item's probe keys and leading probe columns.
constexpr size_t kGroupProbeLds = 16 * 1024;

// Join of a range of probe rows against the group images. A work item of
// THREADS * K probe rows looks every row up (LB in lock step), scans the
// rows' match counts in probe-row order in LDS, reserves its output rows with
// one atomic add, then writes them in order: output row o comes from the
// entry e whose range [e_off[e], e_off[e] + count) holds o (binary search),
// build row e_start[e] + (o - e_off[e]) of the grouped columns. Each thread
// writes four consecutive output rows at a time.
template <int THREADS, int K>
__global__ __launch_bounds__(THREADS) void k_group_join(GroupImage g, const uint64_t *__restrict__ keys, ColSet cols,
                                                        uint64_t rows, OutSink out) {
    constexpr int LB = K < 4 ? K : 4;
    constexpr uint32_t N = THREADS * K;
    static_assert(K % LB == 0, "lookup batches");
    __shared__ uint32_t e_off[N];
    __shared__ uint32_t e_start[N];
    __shared__ uint32_t scan_tmp[THREADS / 64 + 1];
    __shared__ unsigned long long base_s;
    __shared__ __align__(16) unsigned char sprobe[kGroupProbeLds];
    const uint64_t sb = static_cast<uint64_t>(blockIdx.x) * N;
    if (sb >= rows)
        return; // uniform
    const uint32_t n_here = static_cast<uint32_t>(min(static_cast<uint64_t>(N), rows - sb));
#pragma unroll
    for (int gq = 0; gq < K / LB; ++gq) {
        uint64_t key[LB];
        bool valid[LB];
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            const uint32_t e = static_cast<uint32_t>((gq * LB + q) * THREADS) + threadIdx.x;
            valid[q] = e < n_here;
            key[q] = keys[sb + min(e, n_here - 1)];
        }
        uint32_t st[LB], cn[LB];
        group_find<LB>(g, key, valid, st, cn);
#pragma unroll
        for (int q = 0; q < LB; ++q) {
            const uint32_t e = static_cast<uint32_t>((gq * LB + q) * THREADS) + threadIdx.x;
            e_start[e] = st[q];
            e_off[e] = cn[q];
        }
    }
    __syncthreads();
    uint32_t loc[K];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        loc[k] = e_off[threadIdx.x * K + k];
        sum += loc[k];
    }
    uint32_t total;
    uint32_t off = block_excl_scan<THREADS>(sum, scan_tmp, &total);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        e_off[threadIdx.x * K + k] = off;
        off += loc[k];
    }
    if (total == 0)
        return; // uniform: the block total
    if (threadIdx.x == 0)
        base_s = atomicAdd(out.cursor, static_cast<unsigned long long>(total));
    // The item's probe keys and the leading probe columns that fit go to
    // LDS, so the emission below reads its probe side without memory loads
    // (a load issued after the previous rows' stores waits for them).
    uint32_t nlds = 0; // probe columns staged
    {
        size_t used = emit_rcol_bytes(N, 8);
        while (nlds < cols.n && used + emit_rcol_bytes(N, type_size(col_type(cols, nlds))) <= kGroupProbeLds)
            used += emit_rcol_bytes(N, type_size(col_type(cols, nlds++)));
        stage_scols<THREADS>(keys, cols, nlds, static_cast<uint32_t>(sb), n_here, N, sprobe);
    }
    __syncthreads();
    const uint64_t base = base_s;
    const bool pack_ok = (reinterpret_cast<uint64_t>(out.keys) & 15) == 0 && cols_pack_aligned(out.left) &&
                         cols_pack_aligned(out.right);
    // Output rows in aligned groups of four consecutive rows per thread: the
    // rows of a group mostly come from one probe row's consecutive build rows
    // (one search per group), and each column of a full group leaves as one
    // packed store.
    const uint32_t lead = static_cast<uint32_t>(base & 3u);
    const uint64_t a0 = base - lead;
    const uint32_t ngroups = (lead + total + 3) / 4;
    for (uint32_t g0 = 0; g0 < ngroups; g0 += THREADS) {
        const uint32_t gi = g0 + threadIdx.x;
        const int64_t o0 = static_cast<int64_t>(gi) * 4 - lead; // item-relative output row of the group's first row
        uint32_t oc = static_cast<uint32_t>(min<int64_t>(max<int64_t>(o0, 0), total - 1));
        uint32_t e = upper_bound_idx(e_off, N, oc); // the probe row whose rows hold oc (empty rows share its start)
        uint32_t sidx[4], bidx[4];
        bool on[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t o = o0 + r;
            on[r] = gi < ngroups && o >= 0 && o < static_cast<int64_t>(total) && base + o < out.cap;
            if (on[r])
                oc = static_cast<uint32_t>(o); // rows without output repeat the last row with one
            const uint32_t next = e + 1 < N ? e_off[e + 1] : total;
            if (oc >= next) // the row belongs to a later probe row
                e = upper_bound_idx(e_off, N, oc);
            sidx[r] = e;
            bidx[r] = e_start[e] + (oc - e_off[e]);
        }
        const uint64_t p = a0 + static_cast<uint64_t>(gi) * 4;
        const bool packed = pack_ok && on[0] && on[1] && on[2] && on[3];
        move4<uint64_t>(reinterpret_cast<const uint64_t *>(sprobe), sidx, out.keys, p, on, packed);
        move4_set(nullptr, 0, 0u, g.cols, 0, out.left, bidx, p, on, packed);
        move4_set(sprobe + emit_rcol_bytes(N, 8), N, nlds, cols, sb, out.right, sidx, p, on, packed);
    }
}

size_t group_image_lds_bytes(uint32_t max_rows, uint32_t load_pct) { return group_lds(max_rows, load_pct); }

void group_images(hipStream_t s, const uint64_t *r_keys, const uint32_t *r_off, uint32_t n_leaves, uint32_t max_rows,
                  uint32_t load_pct, Cell16 *cells, uint64_t stride, uint32_t *side, uint32_t *perm, KernelTimer *tm) {
    if (n_leaves == 0)
        return;
    const size_t lds = group_lds(max_rows, load_pct);
    TimedScope ts(tm, "group_image", s);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_group_image<512>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    k_group_image<512><<<n_leaves, 512, lds, s>>>(r_keys, r_off, max_rows, load_pct, cells, stride, side, perm);
}

void gather_perm(hipStream_t s, const ColSet &in, const OutColSet &out, const uint32_t *perm, uint64_t rows,
                 KernelTimer *tm) {
    if (rows == 0 || in.n == 0)
        return;
    TimedScope ts(tm, "group_image", s);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((rows + 255) / 256, 256 * 64));
    k_gather_perm<<<grid, 256, 0, s>>>(in, out, perm, rows);
}

void group_count(hipStream_t s, const GroupImage &g, const uint64_t *keys, uint64_t rows, unsigned long long *total,
                 KernelTimer *tm) {
    constexpr int T = 256, K = 8;
    if (rows == 0)
        return;
    TimedScope ts(tm, "group_count", s);
    k_group_count<T, K><<<checked_grid((rows + T * K - 1) / (T * K), T, "group_count"), T, 0, s>>>(g, keys, rows, total);
}

void group_join(hipStream_t s, const GroupImage &g, const uint64_t *keys, const ColSet &cols, uint64_t rows,
                const OutSink &out, KernelTimer *tm) {
    constexpr int T = 256, K = 2; // short items: each probe row may expand to many output rows
    if (rows == 0)
        return;
    TimedScope ts(tm, "leaf_join", s);
    k_group_join<T, K><<<checked_grid((rows + T * K - 1) / (T * K), T, "group_join"), T, 0, s>>>(g, keys, cols, rows,
                                                                                                      out);
}

} // namespace phjd
