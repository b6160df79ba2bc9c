/*
 * oracle.h — CPU oracle for the phj-bench hash-join hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests, the smoke
 * check and the CPU-baseline leg of the benchmark may load it; the product
 * library (libphj_hip.so) never links, loads or calls it.
 *
 * What it restates (reference paths relative to the reference repo root):
 *   orc_int_hash64           src/Hash.h:18-26 (intHash64), :33-58 (batch)
 *   orc_reference_join_*     src/Reference.cpp:48-76 (std::unordered_map
 *                            inner join, probe order, chain in build order)
 *   orc_serialize_rows       src/Reference.cpp:22-43, :79-110 (row bytes
 *                            key | build cols | probe cols)
 *   orc_sort_rows            src/Reference.cpp:113-133 (sorted multiset)
 *   orc_gen_*                src/DataGen.cpp:84-139 for the reference
 *                            streams; the other distributions are the
 *                            harness's own closed forms.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   intHash64: closed-form known answers, bijection, AVX/scalar agreement
 *     and the compiled reference (oracle/_ref) where it builds.
 *   reference join: brute-force nested-loop join on small inputs, the
 *     compiled reference's referenceOutput() where it builds, and the
 *     committed golden fixtures in tests/golden/.
 *   generators: distinctness/range invariants, the reference's own
 *     DataGen (via oracle/_ref) for the reference streams.
 *   checksum: permutation invariance, agreement with a pure-Python
 *     restatement, sensitivity to single-byte changes.
 */
#ifndef PHJ_ORACLE_H
#define PHJ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_COLS 8

typedef struct orc_relation {
    uint64_t rows;
    const uint64_t *keys;
    uint32_t n_cols;
    uint32_t reserved;
    uint32_t types[ORC_MAX_COLS]; /* PayloadType codes 0..4 */
    const void *cols[ORC_MAX_COLS];
} orc_relation;

typedef struct orc_gen_spec {
    uint64_t seed;
    uint64_t build_rows; /* distinct keys for the Zipf distribution */
    uint64_t probe_rows;
    uint32_t distribution; /* 0 uniform, 1 reference streams, 2 zipf */
    uint32_t match_ppm;
    uint32_t build_n_cols;
    uint32_t probe_n_cols;
    uint8_t build_types[ORC_MAX_COLS];
    uint8_t probe_types[ORC_MAX_COLS];
    uint32_t zipf_build_max_dup;
    uint32_t zipf_build_s_milli;
    uint32_t zipf_probe_s_milli;
    uint32_t reserved;
} orc_gen_spec;

typedef struct orc_checksum {
    uint64_t rows;
    uint64_t lane0;
    uint64_t lane1;
} orc_checksum;

size_t orc_type_size(uint32_t type);

/* intHash64 / intHash64Batch. */
uint64_t orc_int_hash64(uint64_t x);
void orc_int_hash64_batch(const uint64_t *in, uint64_t n, uint64_t *out);

/* Generators: fill rows [i0, i0+n) of one side into caller arrays.
 * side 0 = build, 1 = probe. Arrays are indexed from 0 (row i0 -> index 0). */
uint64_t orc_gen_build_rows(const orc_gen_spec *spec); /* total build rows (Zipf duplicates included) */
int orc_gen_fill(const orc_gen_spec *spec, int side, uint64_t i0, uint64_t n, uint64_t *keys, void *const *cols,
                 uint32_t threads);

/* Reference join (src/Reference.cpp:48-76). Count, then fill probe/build
 * row index pairs in the reference's emission order. */
uint64_t orc_reference_join_count(const uint64_t *build_keys, uint64_t nb, const uint64_t *probe_keys, uint64_t np);
uint64_t orc_reference_join_pairs(const uint64_t *build_keys, uint64_t nb, const uint64_t *probe_keys, uint64_t np,
                                  uint64_t *probe_idx, uint64_t *build_idx, uint64_t cap);
/* Same join, digesting each emitted row instead of storing it. */
int orc_reference_join_checksum(const orc_relation *build, const orc_relation *probe, orc_checksum *out);

/* Serialise output columns into n x row_size bytes (row-major), then sort
 * rows lexicographically (std::string order for equal lengths). */
uint64_t orc_row_size(const uint32_t *left_types, uint32_t n_left, const uint32_t *right_types, uint32_t n_right);
void orc_serialize_rows(uint64_t n, const uint64_t *keys, const uint32_t *left_types, const void *const *left,
                        uint32_t n_left, const uint32_t *right_types, const void *const *right, uint32_t n_right,
                        uint8_t *out);
void orc_sort_rows(uint8_t *rows, uint64_t n, uint64_t row_size);

/* Row digest and column-wise checksum (same definition as the GPU). */
void orc_row_digest(const uint8_t *row, uint64_t row_size, uint64_t *lane0, uint64_t *lane1);
void orc_checksum_rows(const uint8_t *rows, uint64_t n, uint64_t row_size, orc_checksum *out);
void orc_checksum_columns(uint64_t n, const uint64_t *keys, const uint32_t *left_types, const void *const *left,
                          uint32_t n_left, const uint32_t *right_types, const void *const *right, uint32_t n_right,
                          orc_checksum *out);

/* Checksum of the join output implied by a generator spec, computed from the
 * closed forms (no join): uniform and reference streams pair each matching
 * probe row with the build row it was drawn from; Zipf pairs a probe row that
 * drew distinct key d with every build row of d (the rows whose shuffled
 * slots are d's prefix range, found from a tabulated forward shuffle). */
int orc_gen_expected_checksum(const orc_gen_spec *spec, uint32_t threads, orc_checksum *out);

#ifdef __cplusplus
}
#endif

#endif
