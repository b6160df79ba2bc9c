// kernels.h — host-side launchers of the gfx950 kernels (internal to libphj_hip.so).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "common.h"

namespace phjd {

// A launch whose grid would exceed 2^32 - 1 work-items in x (HIP's limit).
struct LaunchTooLarge : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// Grid of `blocks` workgroups of `threads`, refused on the host when the
// launch would exceed that limit.
inline uint32_t checked_grid(uint64_t blocks, uint32_t threads, const char *what) {
    if (blocks * threads > 0xFFFFFFFFull)
        throw LaunchTooLarge(std::string(what) + ": " + std::to_string(blocks) + " workgroups of " +
                             std::to_string(threads) + " threads exceed 2^32 work-items");
    return static_cast<uint32_t>(blocks);
}

// Optional per-launch timing: events recorded around each named launch on
// the launch stream, resolved after the stream synchronises.
class KernelTimer {
public:
    virtual ~KernelTimer() = default;
    virtual void begin(const char *name, hipStream_t s) = 0;
    virtual void end(hipStream_t s) = 0;
};

struct TimedScope {
    KernelTimer *t;
    hipStream_t s;
    TimedScope(KernelTimer *t_, const char *name, hipStream_t s_) : t(t_), s(s_) {
        if (t)
            t->begin(name, s);
    }
    ~TimedScope() {
        if (t)
            t->end(s);
    }
};

// ---------------------------------------------------------------- hashing
void launch_hash_batch(hipStream_t s, const uint64_t *in, uint64_t n, uint64_t *out, KernelTimer *tm);
void launch_copy(hipStream_t s, void *dst, const void *src, uint64_t bytes, KernelTimer *tm);
// mode 0: copy src -> dst; 1: read src (dst receives at most one word); 2: write dst.
// Whole 16-byte words only (bytes / 16 of them).
void launch_stream(hipStream_t s, int mode, void *dst, const void *src, uint64_t bytes, KernelTimer *tm);
// Layout probes. Random 16-byte reads of a table of `words` 16-byte words (a
// power of two); a copy with cached loads and stores; one radix pass's memory
// pattern over n 32-byte rows (columns of 8, 8 and 16 bytes, one after
// another in `in`) into `parts` partitions, written as the same three columns
// (rows == false) or as one array of 32-byte rows (rows == true). n is a
// multiple of 4,096 and parts a power of two <= 4,096.
void launch_probe_random_read(hipStream_t s, const void *table, uint64_t words, uint64_t reads, uint64_t seed,
                              uint32_t *sink);
void launch_probe_copy(hipStream_t s, void *dst, const void *src, uint64_t bytes);
void launch_probe_pass(hipStream_t s, bool rows, const void *in, void *out, uint64_t n, uint32_t parts);
// The column form into three separately allocated output columns (8-, 8- and
// 16-byte elements, n each).
void launch_probe_pass_cols(hipStream_t s, const void *in, uint64_t *ok, uint64_t *o1, void *o2, uint64_t n,
                            uint32_t parts);
void launch_fill_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t v);
void launch_fill_u64(hipStream_t s, uint64_t *p, uint64_t n, uint64_t v);

// ---------------------------------------------------------------- radix partition
// One partition pass: every input segment [seg_off[s], seg_off[s+1]) is split
// by hash bits [shift, shift+bits) into 2^bits output segments (output segment
// s*2^bits + d). Rows stay inside their segment's range.
// Bits of one device radix pass: up to 512 partitions per pass (a 9-bit
// pass costs about 1.2x an 8-bit one; a second pass costs 2x).
constexpr uint32_t kDeviceMaxPassBits = 9;

struct RadixPassScratch {
    uint32_t *seg_tile_first; // 2 * (n_seg + 1): first tile, then first scan chunk, per segment
    uint32_t *hist;           // radix_max_tiles * 2^bits
    uint32_t *block_sums;     // radix_max_chunks * 2^bits (chunk bases)
};
uint32_t radix_tile_rows();
uint64_t radix_max_tiles(uint64_t rows, uint32_t n_seg);
uint64_t radix_max_chunks(uint64_t rows, uint32_t n_seg);
uint64_t radix_scan_blocks(uint64_t elems);
// Scatter variants (phj_tuning.flags bits share these values).
constexpr uint32_t kRadixNtLoads = 0x4;  // input columns read with nontemporal loads
constexpr uint32_t kRadixNtStores = 0x8; // output columns written with nontemporal stores
constexpr uint32_t kRadixHashEveryPass = 0x10; // every histogram hashes the keys (no digit column between passes)
constexpr uint32_t kRadixAtomicRank = 0x40; // scatter ranks rows by LDS atomics (not stable) instead of wave ballots
// Column type of a radix pass input that is not stored: each row's index in
// the pass input (row0 + i), written out as a u32 column (type 2).
constexpr uint32_t kTypeRowIndex = 5;
// Bit window of a pass. Uniform: every segment uses (shift, bits) and input
// segment s yields output segments s * 2^bits + d. Per segment (sb != null):
// sb[s] = shift_s | bits_s << 8 with bits_s <= bits, and output segment d of
// s is out_first[s] + d for d < 2^bits_s (out_first has n_seg + 1 entries).
struct PassBits {
    const uint16_t *sb;
    const uint32_t *out_first;
    uint32_t shift;
    uint32_t bits;
};
inline PassBits uniform_bits(uint32_t shift, uint32_t bits) { return PassBits{nullptr, nullptr, shift, bits}; }
// Digit columns between two passes of uniform windows: `in` holds this pass's
// digit of every input row (one byte per row, from the previous pass), so the
// histogram reads it instead of the keys; `out` receives the next pass's
// digit (hash bits [next_shift, next_shift + next_bits), next_bits <= 8) of
// every output row. Either may be null.
struct PassDigits {
    const uint8_t *in;
    uint8_t *out;
    uint32_t next_shift, next_bits;
};
void radix_pass(hipStream_t s, const uint64_t *keys_in, const ColSet &cin, uint64_t *keys_out, const OutColSet &cout,
                const uint32_t *seg_off, uint32_t n_seg, uint64_t rows, const PassBits &pb,
                uint32_t *out_seg_off, const RadixPassScratch &scr, uint32_t flags, KernelTimer *tm,
                const PassDigits &dig = PassDigits{nullptr, nullptr, 0, 0});

// Exclusive scan of n uint32 (device count: n = *count_ptr * mul when count_ptr != nullptr).
void scan_u32(hipStream_t s, uint32_t *data, uint64_t max_n, const uint32_t *count_ptr, uint32_t mul,
              uint32_t *block_sums, KernelTimer *tm, const char *name);
void launch_u32_to_u64(hipStream_t s, const uint32_t *in, uint64_t *out, uint64_t n);
void launch_seg_single(hipStream_t s, uint32_t *seg_off, uint32_t rows);

// ---------------------------------------------------------------- joins
// Output sink shared by all probe kernels: rows are appended at an atomic
// cursor; rows at or beyond `cap` are counted but not written.
struct OutSink {
    unsigned long long *cursor;
    uint64_t cap;
    uint64_t *keys;
    OutColSet left;
    OutColSet right;
};

struct LeafJoinArgs {
    const uint64_t *r_keys;
    ColSet r_cols;
    const uint32_t *r_off; // n_leaves + 1
    const uint64_t *s_keys;
    ColSet s_cols;
    const uint32_t *s_off; // n_leaves + 1
    uint32_t n_leaves;
    uint32_t chunk_rows;   // probe rows per work item
    uint32_t lds_max_rows; // leaves with more build rows use the HBM table path
    uint32_t load_pct;
    const uint32_t *work_first; // n_leaves + 1, exclusive scan of work items per leaf
    const uint32_t *item_leaf;  // leaf of each work item
    uint32_t *item_count;       // probe pass: output rows per work item (max_work + 1 entries)
    const uint32_t *item_off;   // emit pass: exclusive scan of item_count
    uint16_t *heads;            // per probe row: the matched leaf-local build row, or its group's start | kMore
    uint16_t *next16;           // per work item, 2 x leaf rows: each key's rows as one group (perm), group lengths (glen)
    const uint32_t *chain_first; // n_leaves + 1: first item slot of each leaf's items (nr slots per item, 2 entries each)
    uint32_t *chain_flag;       // probe pass sets it when any probe row has several matches
    OutSink out;                // emit pass: rows land at out_base + item_off[item] + ...
    uint64_t out_base;
    uint32_t r_lds;             // multi-match emit: nonzero when the leaf's build columns are staged in LDS
    uint32_t s_lds;             // multi-match emit: nonzero when each round's probe keys are staged in LDS,
                                // with the first s_lds - 1 probe columns
    // Batched joins (fused path): leaf -> join, and per join an output cursor,
    // the end of its output region (capacity) and a repeated-key flag. When
    // leaf_join is null the fused join uses out.cursor, out.cap, chain_flag.
    const uint32_t *leaf_join;
    unsigned long long *join_cursor;
    const uint64_t *join_end;
    uint32_t *join_flag;
    // Prebuilt tables (prebuilt_join): leaf L's table of 16-byte cells is at
    // table_img + L * img_stride, its empty-key side head img_side[L].
    const unsigned char *table_img;
    uint64_t img_stride;
    const uint32_t *img_side;
    // Narrow probe side (fused join only): s_cols is empty and s_rowid holds
    // each partitioned probe row's index in the unpartitioned relation, whose
    // columns are s_src; matched rows gather their probe columns from there.
    const uint32_t *s_rowid;
    ColSet s_src;
};
// Work items per leaf for the LDS path (0 for empty or oversize leaves),
// scanned, and the leaf of each item.
void leaf_work_plan(hipStream_t s, const LeafJoinArgs &a, uint32_t *work_first, uint32_t *scan_tmp, uint32_t *item_leaf,
                    KernelTimer *tm);
uint64_t leaf_max_work(uint64_t probe_rows, uint32_t n_leaves, uint32_t chunk_rows);
// chain_first (n_leaves + 1 entries) from the work plan: each work item of a
// leaf exports its own row groups (nr slots of two entries), scanned over leaves.
void leaf_chain_slots(hipStream_t s, const LeafJoinArgs &a, uint32_t *chain_first, uint32_t *scan_tmp);
// Probe pass: LDS tables, heads/next16, item_count. Emit pass: the rows.
void leaf_probe(hipStream_t s, const LeafJoinArgs &a, uint64_t max_work, uint32_t threads, KernelTimer *tm);
void leaf_emit(hipStream_t s, const LeafJoinArgs &a, uint64_t max_work, bool chains, KernelTimer *tm);
// Fused probe + emit for build sides without repeated keys: one output
// reservation per work item on out.cursor (rows at or beyond out.cap are
// counted only); a leaf with a repeated key sets *chain_flag and writes
// nothing. Work items are probe chunks of leaf_join_fused_chunk() rows.
uint32_t leaf_join_fused_chunk();
void leaf_join_fused(hipStream_t s, const LeafJoinArgs &a, uint64_t max_work, KernelTimer *tm);
// Bytes of a prebuilt leaf table (16-byte cells, the fused join's capacity)
// for a leaf of `rows` build rows, and the tables of every non-empty leaf
// built once into img + L * stride (stride >= leaf_table_bytes(lds_max_rows))
// with the side heads in side[L]; *dup (zeroed by the caller) is set when a
// leaf repeats a key. For joins that probe the same build side many times
// with few probe rows per leaf.
size_t leaf_table_bytes(uint32_t rows, uint32_t load_pct);
// Join of probe_rows partitioned probe rows (a.s_keys, a.s_cols) against the
// prebuilt tables (a.table_img, a.img_stride, a.img_side) of a build side
// with 2^leaf_bits leaves (a.r_off, a.r_cols) and no repeated key; output
// rows are reserved on a.out.cursor (rows at or beyond a.out.cap are counted
// only).
void prebuilt_join(hipStream_t s, const LeafJoinArgs &a, uint64_t probe_rows, uint32_t leaf_bits, KernelTimer *tm);
void leaf_table_images(hipStream_t s, const LeafJoinArgs &a, unsigned char *img, uint64_t stride, uint32_t *side,
                       uint32_t *dup, KernelTimer *tm);

size_t leaf_join_lds_bytes(uint32_t lds_max_rows, uint32_t load_pct);
uint32_t leaf_join_lds_max_rows(uint32_t lds_bytes, uint32_t load_pct);

// Global-memory multi-table hash table with 16-byte (key, RowRefCell) cells
// (src/HashTable.h:35-40): CHJ's 256 sub-tables, oversize PHJ leaves and the
// standalone JoinHashTable entry point. Table t owns cells
// [cell_off[t], cell_off[t] + cap[t]) with cap[t] a power of two.
struct Cell16 {
    uint64_t key;
    uint32_t block_no;
    uint32_t row_no;
};
static_assert(sizeof(Cell16) == 16, "cell is 16 bytes");

// Group images: every leaf's table of a build side with repeated keys, built
// once in HBM as 16-byte cells {key, start, count} (the JoinHashTable cell of
// src/HashTable.h:35-40 with the RowRefCell replaced by the key's row group):
// the key's build rows are rows [start, start + count) of the grouped build
// columns, the partitioned build rows reordered so that the rows of one key
// are adjacent (src/BlockStore.h:56-89 chains, laid out once). The group of
// the empty-key sentinel of leaf L is {side[2L], side[2L + 1]}.
struct GroupImage {
    const Cell16 *cells; // leaf L's table: cells + L * stride, capacity table_cap(rows of L)
    uint64_t stride;     // cells per leaf
    const uint32_t *side;
    const uint32_t *r_off; // leaf offsets of the partitioned build side (n_leaves + 1)
    ColSet cols;           // grouped build columns
    uint32_t leaf_bits;
    uint32_t load_pct;
};
// LDS bytes of the image build for leaves of up to max_rows rows.
size_t group_image_lds_bytes(uint32_t max_rows, uint32_t load_pct);
// One workgroup per leaf (leaves of at most max_rows rows): cells, side
// groups, and perm (grouped row -> partitioned build row).
void group_images(hipStream_t s, const uint64_t *r_keys, const uint32_t *r_off, uint32_t n_leaves, uint32_t max_rows,
                  uint32_t load_pct, Cell16 *cells, uint64_t stride, uint32_t *side, uint32_t *perm, KernelTimer *tm);
// Output row i of every column = input row perm[i], for i < rows.
void gather_perm(hipStream_t s, const ColSet &in, const OutColSet &out, const uint32_t *perm, uint64_t rows,
                 KernelTimer *tm);
// Output rows of `rows` probe keys against the images, added to *total.
void group_count(hipStream_t s, const GroupImage &g, const uint64_t *keys, uint64_t rows, unsigned long long *total,
                 KernelTimer *tm);
// Join of probe rows (keys, cols) against the images: each work item of
// 512 probe rows reserves its output rows on out.cursor (rows at or
// beyond out.cap are counted only) and writes them with the key's rows
// expanded from the grouped build columns.
void group_join(hipStream_t s, const GroupImage &g, const uint64_t *keys, const ColSet &cols, uint64_t rows,
                const OutSink &out, KernelTimer *tm);

struct GlobalHT {
    Cell16 *cells;            // cell i at byte (i << cell_shift): a Cell16 header, then payload
    const uint64_t *cell_off; // n_tables + 1
    uint32_t *next;           // chain successor per build row (row-index space)
    uint32_t *side_head;      // per table: head row of key == kEmptyKey
    uint32_t n_tables;
    uint32_t block_rows;      // RowRefCell{row / block_rows, row % block_rows}
    // kHtLeaves: RowRefCell{leaf, row - build_seg_off[leaf]}
    const uint32_t *build_seg_off;
    // Cells wider than 16 bytes (cell_shift 5 or 6; 0 means 16-byte cells)
    // also hold the build columns `pay` of a key's first row after the header,
    // at ht_pay_next offsets; they are read only for keys with one build row.
    uint32_t cell_shift;
    ColSet pay;
    // Grouped chains (CHJ): when `grouped`, the rows of every key with several
    // build rows are laid out once, after the build, as one contiguous group
    // of `grp` (the build columns, group by group); such a key's cell then
    // holds block_no = its group's first row and row_no = its row count with
    // the several-rows flag. dups is set (to 1) by any insert that displaces
    // a previous head.
    uint32_t grouped;
    ColSet grp;
    uint32_t *dups;
};

// Offset of the next inline payload column of width w after byte `off` of a
// cell (16-byte columns 16-aligned, others aligned to their width).
__host__ __device__ inline uint32_t ht_pay_at(uint32_t off, uint32_t w) {
    const uint32_t a = w >= 16 ? 16u : w;
    return (off + a - 1) & ~(a - 1);
}
// log2 of the cell size that holds a 16-byte header and payload columns of
// these widths inline, or 0 when they need more than 64 bytes.
inline uint32_t ht_wide_shift(const ColSet &cs) {
    if (cs.n == 0)
        return 0;
    uint32_t off = 16;
    for (uint32_t c = 0; c < cs.n; ++c)
        off = ht_pay_at(off, type_size(cs.types[c])) + type_size(cs.types[c]);
    return off <= 32 ? 5u : off <= 64 ? 6u : 0u;
}

// Largest segment (rows) of an offsets array, atomically maxed into *out.
void leaf_max_rows(hipStream_t s, const uint32_t *off, uint32_t n, uint32_t *out);
// Leaves with more than lds_max build rows and a non-empty probe side.
void leaf_big_list(hipStream_t s, const uint32_t *r_off, const uint32_t *s_off, uint32_t n, uint32_t lds_max,
                   uint32_t *leaf_of, uint32_t *count);

// Table selection: CHJ picks table = hash & (n_tables - 1) and homes on
// hash >> 8; the leaf mode maps contiguous row ranges (leaf segments) to tables
// and homes on the low hash bits.
enum HtMode : uint32_t { kHtSingle = 0, kHtChj = 1, kHtLeaves = 2 };

struct HtRows {
    const uint64_t *keys;
    uint64_t rows;
    // kHtLeaves: table b covers rows [seg_off[leaf_of[b]], seg_off[leaf_of[b]+1]);
    // first[b] is the exclusive scan of those row counts.
    const uint32_t *seg_off;
    const uint32_t *leaf_of;
    const uint32_t *first;
};

void ht_chj_plan(hipStream_t s, const uint64_t *keys, uint64_t rows, uint32_t n_tables, uint32_t *counts, KernelTimer *tm);
void ht_init(hipStream_t s, const GlobalHT &ht, uint64_t total_cells, uint64_t build_rows, KernelTimer *tm);
void ht_build(hipStream_t s, const GlobalHT &ht, HtMode mode, const HtRows &r, uint64_t max_rows, KernelTimer *tm,
              const char *name);
// Probe and emit (keys + build cols + probe cols) into `out`.
void ht_probe_emit(hipStream_t s, const GlobalHT &ht, HtMode mode, const HtRows &probe, uint64_t max_rows,
                   const ColSet &r_cols, const ColSet &s_cols, const OutSink &out, KernelTimer *tm, const char *name);
// Rows of every flagged cell's chain (0 for other cells), cnt[total_cells] = 0.
void ht_group_count(hipStream_t s, const GlobalHT &ht, HtMode mode, uint64_t total_cells, uint32_t *cnt, KernelTimer *tm);
// Lists each flagged cell's chain rows (build row indices) at goff[cell] of
// rows[] and turns the cell into {group start, row count | flag}; the build
// columns are then gathered into ht.grp through rows[] (gather_perm).
void ht_group_fill(hipStream_t s, const GlobalHT &ht, HtMode mode, uint64_t total_cells, const uint32_t *goff,
                   uint32_t *rows, KernelTimer *tm);
// Probe only: heads per probe key (RowRefCell), plus next chain export.
void ht_find(hipStream_t s, const GlobalHT &ht, const uint64_t *keys, uint64_t n, phjd::Cell16 *unused, uint32_t *heads_rows,
             KernelTimer *tm);

// ---------------------------------------------------------------- row gathers
// One source relation of gather_segments (device array of these).
struct GatherSource {
    const uint64_t *keys;
    const void *cols[kMaxCols];
};
// Output rows [seg_start[g], seg_start[g+1]) (n_seg + 1 entries, device) are
// rows seg_row[g] .. of srcs[seg_src[g]] (n_src sources, every segment
// non-empty); out's column types are the sources'.
void gather_segments(hipStream_t s, const uint64_t *seg_start, uint32_t n_seg, const uint32_t *seg_src,
                     const uint64_t *seg_row, const GatherSource *srcs, uint32_t n_src, uint64_t total_rows,
                     uint64_t *out_keys, const OutColSet &out, KernelTimer *tm);

// ---------------------------------------------------------------- checksum / generators
void checksum_columns(hipStream_t s, uint64_t n, const uint64_t *keys, const ColSet &left, const ColSet &right,
                      unsigned long long *acc3, KernelTimer *tm);

} // namespace phjd
