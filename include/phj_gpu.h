/*
 * phj_gpu.h — C ABI of the MI355X (gfx950) hash-join library libphj_hip.so.
 *
 * The entry points replace the scheme functions and hot-path primitives of
 * phj-bench (reference paths relative to the reference repository root):
 *
 *   phj_run_chj          ChjResult runCHJ(build, probe, threads)              src/CHJ.h:31
 *   phj_run_phj_pure     PhjResult runPHJ(build, probe, cfg, threads)         src/PHJ.h:31
 *   phj_run_phj          PhjBepResult runPhjBep(build, probe, cfg, threads,   src/PHJBep.h:97-98
 *                                              bep_budget_mib)
 *   phj_join             scheme dispatch by SchemeChoice                      src/Main.cpp:54-109, src/CLI.h:15-22
 *   phj_int_hash64_batch intHash64Batch(in, n, out)                           src/Hash.h:33-58
 *   phj_radix_shuffle    radixShuffle(input, cfg, threads)                    src/RadixPartition.h:195-199
 *   phj_radix_shuffle_project radixShuffle(input, out_schema, in_col_index,  src/RadixPartition.h:187-193
 *                        cfg, threads)
 *   phj_refine_to_leaves refineToLeaves(input, schema, pass_bits, leaves)     src/RadixPartition.h:393-399
 *   phj_hashtable_build_find  JoinHashTable::reserve/insert + batchFind,      src/HashTable.h:72-138,
 *                        with BlockStore next-chain threading                 src/JoinOps.h:205-228, src/BlockStore.h:86-89
 *
 * Conventions (no C++ types, no exceptions cross this boundary):
 *   - Every function returns a phj_status; PHJ_OK is 0. On error,
 *     phj_last_error(ctx) holds a message. PHJ_ERR_INVALID_ARGUMENT maps the
 *     reference's std::invalid_argument (bad radix config, schema mismatch:
 *     src/RadixPartition.cpp:425-426,554-560, src/CLI.cpp:126,350-355) and
 *     PHJ_ERR_OUT_OF_MEMORY maps std::bad_alloc.
 *   - Relations are column-major: a uint64 key column plus n_cols payload
 *     columns of PayloadType widths 1/2/4/8/16 bytes (src/Types.h:27-52,
 *     src/Block.h:21-63). on_device selects host or device pointers.
 *   - Results live in device memory owned by the context until released.
 *   - threads == 0 is accepted and treated as 1 (src/PHJ.cpp:21-22); the GPU
 *     path does not use it.
 *   - One context = one HIP stream; calls on distinct contexts may run
 *     concurrently; calls on one context must not overlap.
 */
#ifndef PHJ_GPU_H
#define PHJ_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHJ_ABI_VERSION 1u
#define PHJ_MAX_COLS 8u
#define PHJ_MAX_PASSES 8u
/* src/Types.h:14 — INVALID_BLOCK marks an empty RowRefCell. */
#define PHJ_INVALID_BLOCK 0xFFFFFFFFu
/* src/Types.h:108 — pipeline block size; output chunking granularity. */
#define PHJ_PIPELINE_BLOCK_ROWS 10000u

typedef enum phj_status {
    PHJ_OK = 0,
    PHJ_ERR_INVALID_ARGUMENT = 1,
    PHJ_ERR_OUT_OF_MEMORY = 2,
    PHJ_ERR_HIP = 3,
    PHJ_ERR_NO_DEVICE = 4,
    PHJ_ERR_UNSUPPORTED = 5,
    PHJ_ERR_CAPACITY = 6
} phj_status;

/* src/Types.h:27-34 — PayloadType codes. */
typedef enum phj_payload_type {
    PHJ_U8 = 0,
    PHJ_U16 = 1,
    PHJ_U32 = 2,
    PHJ_U64 = 3,
    PHJ_U128 = 4
} phj_payload_type;

/* src/CLI.h:15-22 — SchemeChoice codes for the three schemes. */
typedef enum phj_scheme {
    PHJ_SCHEME_CHJ = 1,
    PHJ_SCHEME_PHJ = 2,      /* best-effort partitioning (runPhjBep) */
    PHJ_SCHEME_PHJ_PURE = 3  /* radix-partitioned (runPHJ) */
} phj_scheme;

/* src/Types.h:92-99 — RowRefCell{block_no, row_no}, 8 bytes. */
typedef struct phj_row_ref {
    uint32_t block_no;
    uint32_t row_no;
} phj_row_ref;

/* src/HashTable.h:35-40 — JoinHashTable::Cell, 16 bytes. */
typedef struct phj_ht_cell {
    uint64_t key;
    phj_row_ref ref;
} phj_ht_cell;

/* src/Types.h:17-24 — UInt128{low, high}. */
typedef struct phj_u128 {
    uint64_t low;
    uint64_t high;
} phj_u128;

/* src/Block.h:21-63 — Column{type, data}. */
typedef struct phj_column {
    uint32_t type; /* phj_payload_type */
    uint32_t reserved;
    const void *data;
} phj_column;

/* src/Block.h:159-164 — a BlockStream flattened to contiguous columns. */
typedef struct phj_relation {
    uint64_t rows;
    const uint64_t *keys;
    uint32_t n_cols;
    uint32_t on_device; /* 1: keys/cols are device pointers */
    phj_column cols[PHJ_MAX_COLS];
} phj_relation;

/* Optional overrides; zero in a field selects the built-in choice. */
typedef struct phj_tuning {
    uint32_t tile_rows;     /* radix scatter rows per workgroup tile */
    uint32_t join_threads;  /* leaf-join workgroup size */
    uint32_t ht_load_pct;   /* leaf hash table load factor, percent (default 50) */
    uint32_t lds_leaf_rows; /* max build rows of a leaf kept in an LDS table */
    uint32_t probe_chunk_rows; /* probe rows per leaf-join work item */
    uint32_t flags;         /* PHJ_TUNE_* bits */
} phj_tuning;

#define PHJ_TUNE_FORCE_GLOBAL_HT 0x1u /* leaf tables in HBM even when they fit LDS */
#define PHJ_TUNE_TWO_PHASE_JOIN 0x2u  /* leaf join as a count pass plus an exact-offset emit pass */
#define PHJ_TUNE_CACHED_LOADS 0x4u    /* radix scatter reads its input with cached loads (default: nontemporal) */
#define PHJ_TUNE_NT_STORES 0x8u       /* radix scatter writes with nontemporal stores */
#define PHJ_TUNE_HASH_EVERY_PASS 0x10u /* each radix pass's histogram hashes the keys again instead of reading
                                          the digit byte column the previous pass's scatter wrote */
#define PHJ_TUNE_NARROW_PROBE 0x20u    /* PHJ-PURE partitions the probe side as (key, row index) only and gathers
                                          matched rows' probe columns from the unpartitioned relation */
#define PHJ_TUNE_ATOMIC_RANK 0x40u     /* radix scatter ranks rows with LDS atomics; partitions then keep no input
                                          order (default: stable wave-ballot ranking, rows in input order) */

typedef struct phj_join_desc {
    uint32_t scheme; /* phj_scheme */
    uint32_t threads;
    phj_relation build;
    phj_relation probe;
    uint32_t n_passes;                   /* RadixConfig::pass_bits.size() */
    uint8_t pass_bits[PHJ_MAX_PASSES];   /* RadixConfig::pass_bits, each 1..16 */
    uint64_t bep_budget_bytes;           /* PHJ (best-effort) global probe budget */
    phj_tuning tuning;
} phj_join_desc;

/* src/JoinOutput.h:80-84 — PhaseTiming. */
typedef struct phj_phase_timing {
    double wall_ms;
    double ns_per_row;
} phj_phase_timing;

/* ChjResult / PhjResult / PhjBepResult (src/CHJ.h:12-19, src/PHJ.h:13-22,
 * src/PHJBep.h:13-39) plus the JoinOutput columns (src/JoinOutput.h:18-77)
 * as device arrays: keys, then build-side columns, then probe-side columns. */
typedef struct phj_join_result {
    uint64_t out_rows;
    uint64_t *keys;
    uint32_t n_left;
    uint32_t n_right;
    uint32_t left_types[PHJ_MAX_COLS];
    uint32_t right_types[PHJ_MAX_COLS];
    void *left[PHJ_MAX_COLS];
    void *right[PHJ_MAX_COLS];
    phj_phase_timing build_shuffle;
    phj_phase_timing build;
    phj_phase_timing probe_shuffle;
    phj_phase_timing probe;
    phj_phase_timing eviction_overhead;
    double e2e_wall_ms;
    uint64_t peak_mem_bytes;
    uint64_t bep_budget_mib;
    uint64_t bep_evictions;
    uint64_t bep_refinements;
    uint64_t bep_peak_bytes;
    uint64_t bep_build_skip_retries;
    uint64_t handle; /* nonzero while the result owns device memory */
} phj_join_result;

/* Output of a radix shuffle: one contiguous column per field, rows grouped
 * by partition; offsets[p]..offsets[p+1] is partition p (device array of
 * parts+1 uint64). Replaces PartitionedShuffleOutput's OutBlock chains
 * (src/RadixPartition.h:151-174). */
typedef struct phj_partitioned {
    uint64_t rows;
    uint64_t parts;
    uint64_t *keys;
    uint64_t *offsets;
    uint32_t n_cols;
    uint32_t col_types[PHJ_MAX_COLS];
    void *cols[PHJ_MAX_COLS];
    uint64_t handle;
} phj_partitioned;

/* Order-independent checksum of a join output: per row a 2 x 64-bit digest
 * of the serialised row bytes key | build cols | probe cols
 * (src/Reference.cpp:22-43); lane sums taken mod 2^64. */
typedef struct phj_checksum {
    uint64_t rows;
    uint64_t lane0;
    uint64_t lane1;
} phj_checksum;

typedef struct phj_ctx phj_ctx;

/* ---- context ---- */
uint32_t phj_abi_version(void);
phj_status phj_ctx_create(int device, phj_ctx **out);
phj_status phj_ctx_destroy(phj_ctx *ctx);
const char *phj_last_error(const phj_ctx *ctx);
phj_status phj_synchronize(phj_ctx *ctx);

/* ---- device memory helpers (stream-ordered on the context stream) ---- */
phj_status phj_dev_alloc(phj_ctx *ctx, uint64_t bytes, void **out);
phj_status phj_dev_free(phj_ctx *ctx, void *ptr);
phj_status phj_memcpy_h2d(phj_ctx *ctx, void *dst, const void *src, uint64_t bytes);
phj_status phj_memcpy_d2h(phj_ctx *ctx, void *dst, const void *src, uint64_t bytes);
phj_status phj_mem_stats(phj_ctx *ctx, uint64_t *current_bytes, uint64_t *peak_bytes);
phj_status phj_reset_peak(phj_ctx *ctx);
/* Scatter placement watch (no reference counterpart): radix passes over at
 * least 2^26 rows are timed, and a pass more than 15% slower than the fastest
 * of its shape marks its largest output block slow in the context's pool;
 * later pass outputs are placed clear of marked blocks where the pool already
 * holds room for them (no device allocation is made for that; other
 * allocations use marked blocks as usual). Reports how many blocks were marked
 * and the bytes currently marked (a marked range is forgotten when its memory
 * goes back to the device). Either pointer may be null. */
phj_status phj_placement_stats(phj_ctx *ctx, uint64_t *marked_blocks, uint64_t *marked_bytes);
/* Upper bound on the device bytes the context's live allocations may hold
 * (0, the default: none). An allocation beyond it fails with
 * PHJ_ERR_OUT_OF_MEMORY, like std::bad_alloc in the reference; a call that
 * fails this way returns every block it held. No reference counterpart: the
 * reference tracks peak memory (src/MemTracker.h) but never caps it. */
phj_status phj_set_mem_limit(phj_ctx *ctx, uint64_t bytes);

/* ---- hash: src/Hash.h:18-58 ---- */
phj_status phj_int_hash64_batch(phj_ctx *ctx, const uint64_t *in, uint64_t n, uint64_t *out, int on_device);

/* ---- radix partition: src/RadixPartition.h:187-199, :393-399 ---- */
phj_status phj_radix_shuffle(phj_ctx *ctx, const phj_relation *input, const uint8_t *pass_bits, uint32_t n_passes,
                             phj_partitioned *out);
/* radixShuffle(input, out_schema, in_col_index, cfg, threads),
 * src/RadixPartition.h:187-193: output column c is input column
 * in_col_index[c], of type out_types[c] (which must equal that column's type;
 * an index past the input's columns or a type that differs is
 * PHJ_ERR_INVALID_ARGUMENT). Columns may be dropped, reordered or repeated. */
phj_status phj_radix_shuffle_project(phj_ctx *ctx, const phj_relation *input, const uint32_t *out_types,
                                     const uint32_t *in_col_index, uint32_t n_out_cols, const uint8_t *pass_bits,
                                     uint32_t n_passes, phj_partitioned *out);
/* Restripe a shuffle made with pass_bits[0] only into leaves through passes
 * 1..n_passes-1 (refine_to_leaves mode); out has prod(2^pass_bits) parts. */
phj_status phj_refine_to_leaves(phj_ctx *ctx, const phj_partitioned *pass1, const uint8_t *pass_bits, uint32_t n_passes,
                                phj_partitioned *out);
phj_status phj_partitioned_release(phj_ctx *ctx, phj_partitioned *p);

/* ---- hash table: src/HashTable.h:72-138, src/JoinOps.h:205-228 ----
 * Builds one table over build_keys (row i has RowRefCell{i / block_rows,
 * i % block_rows}), threads multi-match chains, and looks up probe_keys.
 * heads[j] is the chain head for probe key j (block_no == PHJ_INVALID_BLOCK
 * when absent); next[i] is the chain successor of build row i. */
phj_status phj_hashtable_build_find(phj_ctx *ctx, const uint64_t *build_keys, uint64_t n_build, const uint64_t *probe_keys,
                                    uint64_t n_probe, uint32_t block_rows, int on_device, phj_row_ref *heads,
                                    phj_row_ref *next);

/* JoinHashTable build + find that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * results as phj_hashtable_build_find with on_device = 1 (every pointer is a
 * device pointer, heads/next written in device memory). Its scratch (hash
 * cells, chain successors, one chunk of probe rows) is a fixed 40 MiB array
 * that each GPU holds for the life of the process: at most 2^20 build rows
 * (more fail with PHJ_ERR_CAPACITY); probe rows are unlimited (looked up in
 * chunks of 2^20). All calls on one GPU share that array, from any context:
 * calls that overlap on one GPU race and must be serialised by the caller. */
phj_status phj_hashtable_build_find_noalloc(phj_ctx *ctx, const uint64_t *build_keys, uint64_t n_build,
                                            const uint64_t *probe_keys, uint64_t n_probe, uint32_t block_rows,
                                            phj_row_ref *heads, phj_row_ref *next);

/* ---- schemes ---- */
phj_status phj_join(phj_ctx *ctx, const phj_join_desc *desc, phj_join_result *result);
phj_status phj_run_chj(phj_ctx *ctx, const phj_relation *build, const phj_relation *probe, uint32_t threads,
                       phj_join_result *result);
phj_status phj_run_phj_pure(phj_ctx *ctx, const phj_relation *build, const phj_relation *probe, const uint8_t *pass_bits,
                            uint32_t n_passes, uint32_t threads, phj_join_result *result);
phj_status phj_run_phj(phj_ctx *ctx, const phj_relation *build, const phj_relation *probe, const uint8_t *pass_bits,
                       uint32_t n_passes, uint32_t threads, uint64_t bep_budget_mib, phj_join_result *result);
/* Resident build sides: a build relation radix-partitioned once and kept in
 * HBM, probed any number of times (PHJ-PURE's probe half, src/PHJ.cpp:19-150
 * after the build shuffle). The reference re-shuffles the build stream on
 * every call (README.md:146-151); this handle is what a serving loop keeps
 * between requests. phj_probe_registered may run from any context on the same
 * device, concurrently with other probes of the same handle. Release goes
 * through the registering context; it fails with PHJ_ERR_INVALID_ARGUMENT
 * while any probe of the handle is running (counted on the host). */
typedef struct phj_resident phj_resident;
phj_status phj_build_register(phj_ctx *ctx, const phj_relation *build, const uint8_t *pass_bits, uint32_t n_passes,
                              const phj_tuning *tuning, phj_resident **out);
phj_status phj_probe_registered(phj_ctx *ctx, const phj_resident *build, const phj_relation *probe,
                                const phj_tuning *tuning, phj_join_result *result);
phj_status phj_build_release(phj_ctx *ctx, phj_resident *build);
/* Device bytes held by a resident build side. */
phj_status phj_build_bytes(const phj_resident *build, uint64_t *bytes);

/* Batched PHJ-PURE joins: no counterpart in the reference, which joins one
 * pair per call (runPHJ, src/PHJ.h:31); each results[i] equals what
 * phj_join(descs[i]) returns (same rows; phase times are those of the group
 * the join ran in). Joins whose relations lie back to back in device memory
 * (same schema) are joined together in groups; the results of a group share
 * one output allocation, freed with the last phj_result_release. Every desc
 * must be PHJ_SCHEME_PHJ_PURE with a valid radix configuration. */
phj_status phj_join_batch(phj_ctx *ctx, uint32_t n, const phj_join_desc *descs, phj_join_result *results);
/* Batch tuning: no reference counterpart. Joins of a batch aim for leaves of
 * a number of build rows set per size class, floor(log2(build rows)), in
 * PHJ_BATCH_CLASSES entries (2,048 rows for every class until tuned). The
 * autotuner times, for each size class present in descs, a batch of that
 * class's joins with each candidate leaf size (512, 1,024, 2,048, 4,096;
 * best of two timed runs after a warm-up), keeps the fastest in the context
 * for its later phj_join_batch calls, and returns the table and each class's
 * best time in ms (either output may be null). Outputs of the timing runs
 * are released. */
#define PHJ_BATCH_CLASSES 33u
phj_status phj_batch_autotune(phj_ctx *ctx, uint32_t n, const phj_join_desc *descs, uint32_t *leaf_rows_out,
                              double *class_ms_out);
phj_status phj_batch_set_leaf_rows(phj_ctx *ctx, const uint32_t *leaf_rows);
phj_status phj_result_release(phj_ctx *ctx, phj_join_result *result);
/* Copy output columns to host arrays sized out_rows * width. */
phj_status phj_result_copy_to_host(phj_ctx *ctx, const phj_join_result *result, uint64_t *keys, void *const *left,
                                   void *const *right);
phj_status phj_result_checksum(phj_ctx *ctx, const phj_join_result *result, phj_checksum *out);

/* ---- synthetic workloads (harness; no reference counterpart besides
 * src/DataGen.cpp:84-139 for the reference-stream distribution) ---- */
typedef enum phj_distribution {
    PHJ_DIST_UNIFORM = 0,   /* distinct build keys in [1,2^63); probe matches at match_ppm, misses in [2^63,2^64) */
    PHJ_DIST_REFERENCE = 1, /* the reference DataGen streams (100% match) */
    PHJ_DIST_ZIPF = 2       /* duplicated build keys, Zipf-ranked probe keys */
} phj_distribution;

typedef struct phj_gen_spec {
    uint64_t seed;
    uint64_t build_rows;    /* for PHJ_DIST_ZIPF: number of distinct build keys */
    uint64_t probe_rows;
    uint32_t distribution;  /* phj_distribution */
    uint32_t match_ppm;     /* probe match rate in parts per million */
    uint32_t build_n_cols;
    uint32_t probe_n_cols;
    uint8_t build_types[PHJ_MAX_COLS];
    uint8_t probe_types[PHJ_MAX_COLS];
    uint32_t zipf_build_max_dup;    /* PHJ_DIST_ZIPF: duplicates per key drawn from Zipf on [1,max] */
    uint32_t zipf_build_s_milli;    /* Zipf exponent x1000 for duplicate counts */
    uint32_t zipf_probe_s_milli;    /* Zipf exponent x1000 for probe key ranks */
    uint32_t reserved;
} phj_gen_spec;

/* Allocate and fill device relations; release with phj_relation_release. */
phj_status phj_gen_relations(phj_ctx *ctx, const phj_gen_spec *spec, phj_relation *build, phj_relation *probe);
/* n joins generated back to back into one build and one probe relation
 * (same schema): join i's rows are [build_off[i], build_off[i+1]) and
 * [probe_off[i], probe_off[i+1]) (offset arrays of n + 1 entries). */
phj_status phj_gen_batch(phj_ctx *ctx, uint32_t n, const phj_gen_spec *specs, phj_relation *build, phj_relation *probe,
                         uint64_t *build_off, uint64_t *probe_off);
phj_status phj_relation_release(phj_ctx *ctx, phj_relation *rel);
/* Checksum of the exact join output implied by the generator, computed
 * without joining: each matching probe row pairs with the build row it was
 * drawn from (PHJ_DIST_UNIFORM, PHJ_DIST_REFERENCE), or with every build row
 * of the distinct key it drew (PHJ_DIST_ZIPF). */
phj_status phj_gen_expected_checksum(phj_ctx *ctx, const phj_gen_spec *spec, phj_checksum *out);

/* ---- measurement helpers ---- */
/* Record HIP events around every kernel launch on the context stream (off by
 * default). phj_kernel_times then reports, per kernel name, the summed
 * duration and launch count since phj_kernel_times_reset. */
phj_status phj_set_kernel_timing(phj_ctx *ctx, int enabled);
phj_status phj_kernel_times_reset(phj_ctx *ctx);
phj_status phj_kernel_times(phj_ctx *ctx, const char *kernel, double *total_ms, uint64_t *launches);
/* Names of the kernels timed so far, comma separated, into buf (truncated). */
phj_status phj_kernel_names(phj_ctx *ctx, char *buf, uint64_t cap);
/* Plain device copy of `bytes` (src to dst) timed with HIP events; used to
 * calibrate HBM traffic counters. */
phj_status phj_copy_bench(phj_ctx *ctx, void *dst, const void *src, uint64_t bytes, int reps, double *avg_ms);
/* Harness: the part's streaming rates, the measured ceiling beside the 8 TB/s
 * peak. PHJ_STREAM_COPY copies `bytes` from src to dst, PHJ_STREAM_READ reads
 * src (nontemporal 16-byte loads; dst receives at most one word),
 * PHJ_STREAM_WRITE writes dst (whole 16-byte words: bytes / 16 of them);
 * *avg_ms is the mean of `reps` timed passes after one untimed pass. No
 * reference counterpart. */
#define PHJ_STREAM_COPY 0u
#define PHJ_STREAM_READ 1u
#define PHJ_STREAM_WRITE 2u
phj_status phj_stream_bench(phj_ctx *ctx, void *dst, const void *src, uint64_t bytes, uint32_t mode, int reps,
                            double *avg_ms);
/* Harness: measurements behind the partitioning layout, each on buffers the
 * call allocates from the context's pool and frees again; *avg_ms is the mean
 * of `reps` timed passes after one untimed pass. No reference counterpart.
 *   PHJ_PROBE_RANDOM_READ: `param` random 16-byte reads (0: 2^28) of a table
 *     of `bytes` (rounded down to a power of two, at least 4 KiB): the lookup
 *     rate of a hash table of that size;
 *   PHJ_PROBE_SCRATCH_COPY: `bytes` copied from one buffer to another through
 *     a scratch buffer of `param` bytes, a chunk at a time (written to the
 *     scratch, then read back from it): whether a buffer reused at once is
 *     served by the caches;
 *   PHJ_PROBE_PASS_COLUMNS / PHJ_PROBE_PASS_ROWS: one radix pass's memory
 *     pattern over bytes / 32 rows of 32 bytes (columns of 8, 8 and 16 bytes)
 *     into `param` partitions (a power of two <= 4,096; 4,096-row tiles send
 *     runs of 4096 / param rows), written as three columns or as one array of
 *     32-byte rows. */
#define PHJ_PROBE_RANDOM_READ 0u
#define PHJ_PROBE_SCRATCH_COPY 1u
#define PHJ_PROBE_PASS_COLUMNS 2u
#define PHJ_PROBE_PASS_ROWS 3u
phj_status phj_layout_probe(phj_ctx *ctx, uint32_t mode, uint64_t bytes, uint64_t param, int reps, double *avg_ms);
/* Harness: the dependence of one radix pass on where its buffers land. Writes
 * the PHJ_PROBE_PASS_COLUMNS pattern (256 partitions) over `rows` rows (rounded
 * down to a multiple of 4,096) from one input into each of `sets` (1..8) sets
 * of three output columns (8-, 8- and 16-byte elements). With gap ==
 * PHJ_PLACEMENT_SEPARATE every column is its own fresh hipMalloc; otherwise a
 * set is one fresh hipMalloc holding the columns one after another with `gap`
 * bytes (rounded up to 256) between them. Allocated set by set outside the
 * context's pool and freed on return; ms[k] is the mean of `reps` timed passes
 * into set k after one untimed pass. No reference counterpart. */
#define PHJ_PLACEMENT_SEPARATE UINT64_MAX
phj_status phj_placement_probe(phj_ctx *ctx, uint64_t rows, uint32_t sets, uint64_t gap, int reps, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* PHJ_GPU_H */
