"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and of the compiled
reference (oracle/_ref/libphjref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg. The product (phj-bench_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAX_COLS = 8
TYPE_SIZE = {0: 1, 1: 2, 2: 4, 3: 8, 4: 16}
TYPE_NAME = {0: "u8", 1: "u16", 2: "u32", 3: "u64", 4: "u128"}
NAME_TYPE = {v: k for k, v in TYPE_NAME.items()}
NP_DTYPE = {0: np.uint8, 1: np.uint16, 2: np.uint32, 3: np.uint64}


class OrcRelation(C.Structure):
    _fields_ = [
        ("rows", C.c_uint64),
        ("keys", C.c_void_p),
        ("n_cols", C.c_uint32),
        ("reserved", C.c_uint32),
        ("types", C.c_uint32 * MAX_COLS),
        ("cols", C.c_void_p * MAX_COLS),
    ]


class OrcGenSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("build_rows", C.c_uint64),
        ("probe_rows", C.c_uint64),
        ("distribution", C.c_uint32),
        ("match_ppm", C.c_uint32),
        ("build_n_cols", C.c_uint32),
        ("probe_n_cols", C.c_uint32),
        ("build_types", C.c_uint8 * MAX_COLS),
        ("probe_types", C.c_uint8 * MAX_COLS),
        ("zipf_build_max_dup", C.c_uint32),
        ("zipf_build_s_milli", C.c_uint32),
        ("zipf_probe_s_milli", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class OrcChecksum(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("lane0", C.c_uint64), ("lane1", C.c_uint64)]

    def tuple(self):
        return (int(self.rows), int(self.lane0), int(self.lane1))


class RefResult(C.Structure):
    _fields_ = [
        ("out_rows", C.c_uint64),
        ("checksum", OrcChecksum),
        ("e2e_wall_ms", C.c_double),
        ("build_shuffle_ms", C.c_double),
        ("build_ms", C.c_double),
        ("probe_shuffle_ms", C.c_double),
        ("probe_ms", C.c_double),
        ("peak_mem_bytes", C.c_uint64),
    ]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
        _lib = C.CDLL(path)
        _lib.orc_int_hash64.restype = C.c_uint64
        _lib.orc_int_hash64.argtypes = [C.c_uint64]
        _lib.orc_gen_build_rows.restype = C.c_uint64
        _lib.orc_reference_join_count.restype = C.c_uint64
        _lib.orc_reference_join_pairs.restype = C.c_uint64
        _lib.orc_row_size.restype = C.c_uint64
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libphjref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libphjref.so")
        if not os.path.exists(path):
            raise RuntimeError("compiled reference oracle/_ref/libphjref.so is not built")
        _ref = C.CDLL(path)
        _ref.ref_reference_rows.restype = C.c_uint64
    return _ref


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- relations
class Relation:
    """Host relation: uint64 keys plus payload columns (numpy arrays).

    A u128 column is an (n, 2) uint64 array of (low, high)."""

    def __init__(self, keys, cols, types):
        self.keys = np.ascontiguousarray(keys, dtype=np.uint64)
        self.cols = [np.ascontiguousarray(c) for c in cols]
        self.types = list(types)
        assert len(self.cols) == len(self.types)

    @property
    def rows(self):
        return int(self.keys.shape[0])

    def orc(self) -> OrcRelation:
        r = OrcRelation()
        r.rows = self.rows
        r.keys = _ptr(self.keys)
        r.n_cols = len(self.cols)
        for i, (c, t) in enumerate(zip(self.cols, self.types)):
            r.types[i] = t
            r.cols[i] = _ptr(c)
        return r

    def take(self, idx):
        return Relation(self.keys[idx], [c[idx] for c in self.cols], self.types)


def empty_column(t, n):
    if t == 4:
        return np.zeros((n, 2), dtype=np.uint64)
    return np.zeros(n, dtype=NP_DTYPE[t])


def make_spec(seed=42, build_rows=1000, probe_rows=4000, distribution=0, match_rate=1.0, build_types=(3,),
              probe_types=(3,), zipf_max_dup=64, zipf_build_s=1.2, zipf_probe_s=0.8) -> OrcGenSpec:
    s = OrcGenSpec()
    s.seed = seed
    s.build_rows = build_rows
    s.probe_rows = probe_rows
    s.distribution = distribution
    s.match_ppm = int(round(match_rate * 1_000_000))
    s.build_n_cols = len(build_types)
    s.probe_n_cols = len(probe_types)
    for i, t in enumerate(build_types):
        s.build_types[i] = t
    for i, t in enumerate(probe_types):
        s.probe_types[i] = t
    s.zipf_build_max_dup = zipf_max_dup
    s.zipf_build_s_milli = int(round(zipf_build_s * 1000))
    s.zipf_probe_s_milli = int(round(zipf_probe_s * 1000))
    return s


def gen_side(spec: OrcGenSpec, side: int, i0: int = 0, n: int | None = None, threads: int = 8) -> Relation:
    L = lib()
    total = L.orc_gen_build_rows(C.byref(spec)) if side == 0 else spec.probe_rows
    if n is None:
        n = int(total) - i0
    types = list(spec.build_types[: spec.build_n_cols]) if side == 0 else list(spec.probe_types[: spec.probe_n_cols])
    keys = np.zeros(n, dtype=np.uint64)
    cols = [empty_column(t, n) for t in types]
    colp = (C.c_void_p * MAX_COLS)(*([c.ctypes.data for c in cols] + [None] * (MAX_COLS - len(cols))))
    L.orc_gen_fill(C.byref(spec), C.c_int(side), C.c_uint64(i0), C.c_uint64(n), _ptr(keys), colp, C.c_uint32(threads))
    return Relation(keys, cols, types)


def gen(spec: OrcGenSpec, threads: int = 8):
    return gen_side(spec, 0, threads=threads), gen_side(spec, 1, threads=threads)


# ---------------------------------------------------------------- hashing
def int_hash64(x):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    out = np.empty_like(x)
    lib().orc_int_hash64_batch(_ptr(x), C.c_uint64(x.size), _ptr(out))
    return out


def int_hash64_py(x: int) -> int:
    """Pure-Python closed form of intHash64 (src/Hash.h:18-26)."""
    M = (1 << 64) - 1
    x ^= 0x4CF5AD432745937F
    x = ((x ^ (x >> 33)) * 0xFF51AFD7ED558CCD) & M
    x = ((x ^ (x >> 33)) * 0xC4CEB9FE1A85EC53) & M
    x ^= x >> 33
    return x


# ---------------------------------------------------------------- radix partition
def leaf_ids(keys, total_bits: int):
    """Partition of every key after `total_bits` bits of radix passes: the top
    bits of intHash64 (src/RadixPartition.cpp:546-577, each pass taking the
    next bits below the ones already consumed)."""
    if total_bits == 0:
        return np.zeros(np.asarray(keys).shape[0], dtype=np.int64)
    return (int_hash64(keys) >> np.uint64(64 - total_bits)).astype(np.int64)


def radix_shuffle(rel: Relation, pass_bits, out_types=None, in_col_index=None):
    """radixShuffle(input, out_schema, in_col_index, cfg, threads)
    (src/RadixPartition.cpp:546-577): returns (partitioned relation, offsets).

    Order: scatterBatch appends each batch's rows to their partition in row
    order (src/RadixPartition.cpp:316-393); pass 1 gives each thread a
    contiguous run of blocks and concatenates the per-thread chains in thread
    order (:144-205), and every later pass re-scatters one input partition's
    chain on one thread (:209-250). Every partition therefore holds its rows in
    input order: a stable sort of the rows by partition id."""
    total = int(sum(pass_bits))
    if in_col_index is None:
        in_col_index = list(range(len(rel.cols)))
        out_types = list(rel.types)
    ids = leaf_ids(rel.keys, total)
    order = np.argsort(ids, kind="stable")
    cols = [rel.cols[i][order] for i in in_col_index]
    off = np.zeros((1 << total) + 1, dtype=np.int64)
    off[1:] = np.cumsum(np.bincount(ids, minlength=1 << total))
    return Relation(rel.keys[order], cols, list(out_types)), off


# ---------------------------------------------------------------- join
def reference_join_pairs(build: Relation, probe: Relation):
    L = lib()
    m = L.orc_reference_join_count(_ptr(build.keys), C.c_uint64(build.rows), _ptr(probe.keys), C.c_uint64(probe.rows))
    pi = np.zeros(max(m, 1), dtype=np.uint64)
    bi = np.zeros(max(m, 1), dtype=np.uint64)
    L.orc_reference_join_pairs(_ptr(build.keys), C.c_uint64(build.rows), _ptr(probe.keys), C.c_uint64(probe.rows),
                               _ptr(pi), _ptr(bi), C.c_uint64(m))
    return pi[:m], bi[:m]


class JoinRows:
    """Join output as columns: keys, build-side columns, probe-side columns."""

    def __init__(self, keys, left, left_types, right, right_types):
        self.keys = np.ascontiguousarray(keys, dtype=np.uint64)
        self.left = [np.ascontiguousarray(c) for c in left]
        self.left_types = list(left_types)
        self.right = [np.ascontiguousarray(c) for c in right]
        self.right_types = list(right_types)

    @property
    def rows(self):
        return int(self.keys.shape[0])


def reference_join(build: Relation, probe: Relation) -> JoinRows:
    """src/Reference.cpp:48-76 restated: output in probe order, chains in build order."""
    pi, bi = reference_join_pairs(build, probe)
    return JoinRows(probe.keys[pi], [c[bi] for c in build.cols], build.types, [c[pi] for c in probe.cols], probe.types)


def _types_arr(ts):
    a = (C.c_uint32 * MAX_COLS)()
    for i, t in enumerate(ts):
        a[i] = t
    return a


def _cols_arr(cs):
    return (C.c_void_p * MAX_COLS)(*([c.ctypes.data for c in cs] + [None] * (MAX_COLS - len(cs))))


def row_size(j: JoinRows) -> int:
    return 8 + sum(TYPE_SIZE[t] for t in j.left_types) + sum(TYPE_SIZE[t] for t in j.right_types)


def serialize_sorted(j: JoinRows) -> np.ndarray:
    """Sorted serialised rows (src/Reference.cpp:22-43, :113-133) as (n, row_size) uint8."""
    L = lib()
    rs = row_size(j)
    buf = np.zeros((max(j.rows, 1), rs), dtype=np.uint8)
    L.orc_serialize_rows(C.c_uint64(j.rows), _ptr(j.keys), _types_arr(j.left_types), _cols_arr(j.left),
                         C.c_uint32(len(j.left)), _types_arr(j.right_types), _cols_arr(j.right),
                         C.c_uint32(len(j.right)), _ptr(buf))
    L.orc_sort_rows(_ptr(buf), C.c_uint64(j.rows), C.c_uint64(rs))
    return buf[: j.rows]


def checksum(j: JoinRows):
    out = OrcChecksum()
    lib().orc_checksum_columns(C.c_uint64(j.rows), _ptr(j.keys), _types_arr(j.left_types), _cols_arr(j.left),
                               C.c_uint32(len(j.left)), _types_arr(j.right_types), _cols_arr(j.right),
                               C.c_uint32(len(j.right)), C.byref(out))
    return out.tuple()


def checksum_rows(rows: np.ndarray):
    out = OrcChecksum()
    n, rs = rows.shape if rows.size else (0, 1)
    lib().orc_checksum_rows(_ptr(np.ascontiguousarray(rows)), C.c_uint64(n), C.c_uint64(rs), C.byref(out))
    return out.tuple()


def reference_join_checksum(build: Relation, probe: Relation):
    out = OrcChecksum()
    lib().orc_reference_join_checksum(C.byref(build.orc()), C.byref(probe.orc()), C.byref(out))
    return out.tuple()


def expected_checksum(spec: OrcGenSpec, threads: int = 8):
    out = OrcChecksum()
    rc = lib().orc_gen_expected_checksum(C.byref(spec), C.c_uint32(threads), C.byref(out))
    if rc != 0:
        raise ValueError("closed-form checksum failed")
    return out.tuple()


def row_digest_py(row: bytes):
    """Pure-Python restatement of the row digest (two SplitMix64 lanes)."""
    M = (1 << 64) - 1

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    n = len(row)
    a = 0x243F6A8885A308D3 ^ n
    c = (0x13198A2E03707344 + n) & M
    padded = row + b"\0" * ((-n) % 8)
    for k in range(0, len(padded), 8):
        w = int.from_bytes(padded[k:k + 8], "little")
        a = mix(a ^ w)
        c = mix((c + w * 0x9E3779B97F4A7C15) & M)
    return a, c


# ---------------------------------------------------------------- compiled reference
def ref_int_hash64(x):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    out = np.empty_like(x)
    ref().ref_int_hash64_batch(_ptr(x), C.c_uint64(x.size), _ptr(out))
    return out


def ref_radix_shuffle(rel: Relation, pass_bits, out_types=None, in_col_index=None, threads=1):
    """The compiled reference's radixShuffle, flattened: (relation, offsets)."""
    if in_col_index is None:
        in_col_index = list(range(len(rel.cols)))
        out_types = list(rel.types)
    total = int(sum(pass_bits))
    if total > 20:
        raise NotImplementedError("this wrapper sizes part_rows for at most 2^20 partitions")
    parts = 1 << total
    keys = np.zeros(max(rel.rows, 1), dtype=np.uint64)
    cols = [empty_column(t, max(rel.rows, 1)) for t in out_types]
    rows = np.zeros(parts, dtype=np.uint64)
    pb = (C.c_uint8 * len(pass_bits))(*pass_bits)
    idx = (C.c_uint64 * max(1, len(in_col_index)))(*in_col_index)
    rc = ref().ref_radix_shuffle(C.byref(rel.orc()), _types_arr(out_types), idx, C.c_uint32(len(out_types)), pb,
                                 C.c_uint32(len(pass_bits)), C.c_uint64(threads), _ptr(keys), _cols_arr(cols),
                                 _ptr(rows))
    if rc != 0:
        raise ValueError(f"reference radixShuffle rejected the configuration (rc={rc})")
    off = np.zeros(parts + 1, dtype=np.int64)
    off[1:] = np.cumsum(rows.astype(np.int64))
    n = rel.rows
    return Relation(keys[:n], [c[:n] for c in cols], list(out_types)), off


def ref_reference_checksum(build: Relation, probe: Relation):
    out = OrcChecksum()
    ref().ref_reference_checksum(C.byref(build.orc()), C.byref(probe.orc()), C.byref(out))
    return out.tuple()


def ref_reference_rows(build: Relation, probe: Relation, row_size_: int, max_rows: int):
    buf = np.zeros((max(max_rows, 1), row_size_), dtype=np.uint8)
    n = ref().ref_reference_rows(C.byref(build.orc()), C.byref(probe.orc()), _ptr(buf), C.c_uint64(buf.nbytes))
    assert n <= max_rows, (n, max_rows)
    return buf[:n]


def ref_run(scheme: int, build: Relation, probe: Relation, pass_bits=(8,), threads=1, bep_budget_mib=512,
            want_checksum=True) -> RefResult:
    res = RefResult()
    pb = (C.c_uint8 * len(pass_bits))(*pass_bits)
    rc = ref().ref_run(C.c_int(scheme), C.byref(build.orc()), C.byref(probe.orc()), pb, C.c_uint32(len(pass_bits)),
                       C.c_uint64(threads), C.c_uint64(bep_budget_mib), C.c_int(1 if want_checksum else 0),
                       C.byref(res))
    if rc != 0:
        raise RuntimeError(f"reference run failed (rc={rc})")
    return res


def ref_generate(build_rows, probe_rows, build_types, probe_types, seed=42):
    bkeys = np.zeros(build_rows, dtype=np.uint64)
    pkeys = np.zeros(probe_rows, dtype=np.uint64)
    bcols = [empty_column(t, build_rows) for t in build_types]
    pcols = [empty_column(t, probe_rows) for t in probe_types]
    ref().ref_generate(C.c_uint64(build_rows), C.c_uint64(probe_rows), _types_arr(build_types),
                       C.c_uint32(len(build_types)), _types_arr(probe_types), C.c_uint32(len(probe_types)),
                       C.c_uint64(seed), _ptr(bkeys), _cols_arr(bcols), _ptr(pkeys), _cols_arr(pcols))
    return Relation(bkeys, bcols, build_types), Relation(pkeys, pcols, probe_types)
