"""Python host mirror of the phj-bench scheme API over libphj_hip.so (gfx950).

The names and argument meanings follow the reference's scheme functions:

    runCHJ(build, probe, threads)                            src/CHJ.h:31
    runPHJ(build, probe, cfg, threads)        (PHJ-PURE)     src/PHJ.h:31
    runPhjBep(build, probe, cfg, threads, bep_budget_mib)    src/PHJBep.h:97-98
    intHash64Batch(keys)                                     src/Hash.h:33-58
    radixShuffle(input, [out_schema, in_col_index,] cfg, threads)
                                                             src/RadixPartition.h:187-199
    refineToLeaves(pass1, pass_bits)                         src/RadixPartition.h:393-399

Errors map as in the reference: invalid radix configurations and schemas
raise ValueError (std::invalid_argument), allocation failures MemoryError
(std::bad_alloc); device/runtime failures raise RuntimeError. The library is
required: importing succeeds without it, but every call raises when the HIP
library is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphj_hip.so")
MAX_COLS = 8
MAX_PASSES = 8
INVALID_BLOCK = 0xFFFFFFFF
PIPELINE_BLOCK_ROWS = 10_000  # src/Types.h:108
BATCH_CLASSES = 33  # batch tuning size classes, floor(log2(build rows))

# src/Types.h:27-34
U8, U16, U32, U64, U128 = 0, 1, 2, 3, 4
TYPE_SIZE = {U8: 1, U16: 2, U32: 4, U64: 8, U128: 16}
TYPE_NAME = {U8: "u8", U16: "u16", U32: "u32", U64: "u64", U128: "u128"}
NP_DTYPE = {U8: np.uint8, U16: np.uint16, U32: np.uint32, U64: np.uint64}

# src/CLI.h:15-22
SCHEME_CHJ, SCHEME_PHJ, SCHEME_PHJ_PURE = 1, 2, 3

DIST_UNIFORM, DIST_REFERENCE, DIST_ZIPF = 0, 1, 2

OK, ERR_INVALID_ARGUMENT, ERR_OUT_OF_MEMORY, ERR_HIP, ERR_NO_DEVICE, ERR_UNSUPPORTED, ERR_CAPACITY = range(7)


class _Column(C.Structure):
    _fields_ = [("type", C.c_uint32), ("reserved", C.c_uint32), ("data", C.c_void_p)]


class _Relation(C.Structure):
    _fields_ = [
        ("rows", C.c_uint64),
        ("keys", C.c_void_p),
        ("n_cols", C.c_uint32),
        ("on_device", C.c_uint32),
        ("cols", _Column * MAX_COLS),
    ]


class _Tuning(C.Structure):
    _fields_ = [
        ("tile_rows", C.c_uint32),
        ("join_threads", C.c_uint32),
        ("ht_load_pct", C.c_uint32),
        ("lds_leaf_rows", C.c_uint32),
        ("probe_chunk_rows", C.c_uint32),
        ("flags", C.c_uint32),
    ]


class _JoinDesc(C.Structure):
    _fields_ = [
        ("scheme", C.c_uint32),
        ("threads", C.c_uint32),
        ("build", _Relation),
        ("probe", _Relation),
        ("n_passes", C.c_uint32),
        ("pass_bits", C.c_uint8 * MAX_PASSES),
        ("bep_budget_bytes", C.c_uint64),
        ("tuning", _Tuning),
    ]


class _Phase(C.Structure):
    _fields_ = [("wall_ms", C.c_double), ("ns_per_row", C.c_double)]


class _JoinResult(C.Structure):
    _fields_ = [
        ("out_rows", C.c_uint64),
        ("keys", C.c_void_p),
        ("n_left", C.c_uint32),
        ("n_right", C.c_uint32),
        ("left_types", C.c_uint32 * MAX_COLS),
        ("right_types", C.c_uint32 * MAX_COLS),
        ("left", C.c_void_p * MAX_COLS),
        ("right", C.c_void_p * MAX_COLS),
        ("build_shuffle", _Phase),
        ("build", _Phase),
        ("probe_shuffle", _Phase),
        ("probe", _Phase),
        ("eviction_overhead", _Phase),
        ("e2e_wall_ms", C.c_double),
        ("peak_mem_bytes", C.c_uint64),
        ("bep_budget_mib", C.c_uint64),
        ("bep_evictions", C.c_uint64),
        ("bep_refinements", C.c_uint64),
        ("bep_peak_bytes", C.c_uint64),
        ("bep_build_skip_retries", C.c_uint64),
        ("handle", C.c_uint64),
    ]


class _Partitioned(C.Structure):
    _fields_ = [
        ("rows", C.c_uint64),
        ("parts", C.c_uint64),
        ("keys", C.c_void_p),
        ("offsets", C.c_void_p),
        ("n_cols", C.c_uint32),
        ("col_types", C.c_uint32 * MAX_COLS),
        ("cols", C.c_void_p * MAX_COLS),
        ("handle", C.c_uint64),
    ]


class _Checksum(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("lane0", C.c_uint64), ("lane1", C.c_uint64)]


class _GenSpec(C.Structure):
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


class _RowRef(C.Structure):
    _fields_ = [("block_no", C.c_uint32), ("row_no", C.c_uint32)]


# Exported symbols with their argument types (the C header's declarations).
_SIGNATURES = {
    "phj_abi_version": ([], C.c_uint32),
    "phj_ctx_create": ([C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "phj_ctx_destroy": ([C.c_void_p], C.c_int),
    "phj_last_error": ([C.c_void_p], C.c_char_p),
    "phj_synchronize": ([C.c_void_p], C.c_int),
    "phj_dev_alloc": ([C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)], C.c_int),
    "phj_dev_free": ([C.c_void_p, C.c_void_p], C.c_int),
    "phj_memcpy_h2d": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64], C.c_int),
    "phj_memcpy_d2h": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64], C.c_int),
    "phj_mem_stats": ([C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
    "phj_placement_stats": ([C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
    "phj_reset_peak": ([C.c_void_p], C.c_int),
    "phj_set_mem_limit": ([C.c_void_p, C.c_uint64], C.c_int),
    "phj_int_hash64_batch": ([C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int], C.c_int),
    "phj_radix_shuffle": ([C.c_void_p, C.POINTER(_Relation), C.c_void_p, C.c_uint32, C.POINTER(_Partitioned)], C.c_int),
    "phj_radix_shuffle_project": ([C.c_void_p, C.POINTER(_Relation), C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                   C.c_uint32, C.POINTER(_Partitioned)], C.c_int),
    "phj_refine_to_leaves": ([C.c_void_p, C.POINTER(_Partitioned), C.c_void_p, C.c_uint32, C.POINTER(_Partitioned)],
                             C.c_int),
    "phj_partitioned_release": ([C.c_void_p, C.POINTER(_Partitioned)], C.c_int),
    "phj_hashtable_build_find": ([C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int,
                                  C.c_void_p, C.c_void_p], C.c_int),
    "phj_hashtable_build_find_noalloc": ([C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_uint32,
                                          C.c_void_p, C.c_void_p], C.c_int),
    "phj_join": ([C.c_void_p, C.POINTER(_JoinDesc), C.POINTER(_JoinResult)], C.c_int),
    "phj_build_register": ([C.c_void_p, C.POINTER(_Relation), C.c_void_p, C.c_uint32, C.POINTER(_Tuning),
                            C.POINTER(C.c_void_p)], C.c_int),
    "phj_probe_registered": ([C.c_void_p, C.c_void_p, C.POINTER(_Relation), C.POINTER(_Tuning), C.POINTER(_JoinResult)],
                             C.c_int),
    "phj_build_release": ([C.c_void_p, C.c_void_p], C.c_int),
    "phj_build_bytes": ([C.c_void_p, C.POINTER(C.c_uint64)], C.c_int),
    "phj_join_batch": ([C.c_void_p, C.c_uint32, C.POINTER(_JoinDesc), C.POINTER(_JoinResult)], C.c_int),
    "phj_batch_autotune": ([C.c_void_p, C.c_uint32, C.POINTER(_JoinDesc), C.c_void_p, C.c_void_p], C.c_int),
    "phj_batch_set_leaf_rows": ([C.c_void_p, C.c_void_p], C.c_int),
    "phj_run_chj": ([C.c_void_p, C.POINTER(_Relation), C.POINTER(_Relation), C.c_uint32, C.POINTER(_JoinResult)], C.c_int),
    "phj_run_phj_pure": ([C.c_void_p, C.POINTER(_Relation), C.POINTER(_Relation), C.c_void_p, C.c_uint32, C.c_uint32,
                          C.POINTER(_JoinResult)], C.c_int),
    "phj_run_phj": ([C.c_void_p, C.POINTER(_Relation), C.POINTER(_Relation), C.c_void_p, C.c_uint32, C.c_uint32,
                     C.c_uint64, C.POINTER(_JoinResult)], C.c_int),
    "phj_result_release": ([C.c_void_p, C.POINTER(_JoinResult)], C.c_int),
    "phj_result_copy_to_host": ([C.c_void_p, C.POINTER(_JoinResult), C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "phj_result_checksum": ([C.c_void_p, C.POINTER(_JoinResult), C.POINTER(_Checksum)], C.c_int),
    "phj_gen_relations": ([C.c_void_p, C.POINTER(_GenSpec), C.POINTER(_Relation), C.POINTER(_Relation)], C.c_int),
    "phj_gen_batch": ([C.c_void_p, C.c_uint32, C.POINTER(_GenSpec), C.POINTER(_Relation), C.POINTER(_Relation),
                       C.c_void_p, C.c_void_p], C.c_int),
    "phj_relation_release": ([C.c_void_p, C.POINTER(_Relation)], C.c_int),
    "phj_gen_expected_checksum": ([C.c_void_p, C.POINTER(_GenSpec), C.POINTER(_Checksum)], C.c_int),
    "phj_set_kernel_timing": ([C.c_void_p, C.c_int], C.c_int),
    "phj_kernel_times_reset": ([C.c_void_p], C.c_int),
    "phj_kernel_times": ([C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)], C.c_int),
    "phj_kernel_names": ([C.c_void_p, C.c_char_p, C.c_uint64], C.c_int),
    "phj_copy_bench": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "phj_stream_bench": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.POINTER(C.c_double)],
                         C.c_int),
    "phj_layout_probe": ([C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "phj_placement_probe": ([C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64, C.c_int, C.POINTER(C.c_double)],
                            C.c_int),
}

_lib = None


def library_path() -> str:
    return LIB_PATH


def load_library():
    """Load libphj_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libphj_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


class PhjError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"phj error {code}: {msg}")
        self.code = code


def _raise(ctx_ptr, rc):
    if rc == OK:
        return
    msg = load_library().phj_last_error(ctx_ptr)
    msg = msg.decode() if msg else ""
    if rc == ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if rc == ERR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    raise PhjError(rc, msg)


# ---------------------------------------------------------------- types
@dataclass
class RadixConfig:
    """src/RadixPartition.h:20-33 — bits per pass, consumed from the top of the hash."""
    pass_bits: list = field(default_factory=lambda: [8])

    def total_bits(self):
        return sum(self.pass_bits)

    def partitions(self):
        return 1 << self.total_bits()

    @staticmethod
    def from_partitions(partitions=256, passes=1):
        """src/CLI.cpp:323-348 — total bits split evenly, remainder to the first passes."""
        total = partitions.bit_length() - 1
        if passes > total:
            raise ValueError("passes cannot exceed log2(partitions)")
        base, extra = divmod(total, passes)
        return RadixConfig([base + (1 if p < extra else 0) for p in range(passes)])

    @staticmethod
    def for_build_rows(rows, leaf_rows=2048):
        """Leaves of at most leaf_rows build rows (at least 1 bit), passes of at most 8 bits."""
        total = 1
        while total < 16 and (rows >> total) > leaf_rows:
            total += 1
        return RadixConfig.from_partitions(1 << total, (total + 7) // 8)


@dataclass
class PhaseTiming:
    wall_ms: float = 0.0
    ns_per_row: float = 0.0


class HostRelation:
    """Column-major host relation: uint64 keys + payload columns (numpy).

    A u128 column is an (n, 2) uint64 array of (low, high) (src/Types.h:17-24)."""

    def __init__(self, keys, cols=(), types=()):
        self.keys = np.ascontiguousarray(keys, dtype=np.uint64)
        self.cols = [np.ascontiguousarray(c) for c in cols]
        self.types = list(types)
        if len(self.cols) != len(self.types):
            raise ValueError("schema and column count mismatch")
        for c, t in zip(self.cols, self.types):
            want = (self.rows, 2) if t == U128 else (self.rows,)
            if c.shape != want or c.dtype != (np.uint64 if t == U128 else NP_DTYPE[t]):
                raise ValueError(f"column of type {TYPE_NAME[t]} has shape {c.shape} dtype {c.dtype}")

    @property
    def rows(self):
        return int(self.keys.shape[0])

    def _c(self) -> _Relation:
        r = _Relation()
        r.rows = self.rows
        r.keys = self.keys.ctypes.data
        r.n_cols = len(self.cols)
        r.on_device = 0
        for i, (c, t) in enumerate(zip(self.cols, self.types)):
            r.cols[i].type = t
            r.cols[i].data = c.ctypes.data
        return r


class DeviceRelation:
    """A relation resident in HBM (owned by the context)."""

    def __init__(self, ctx, crel: _Relation):
        self.ctx = ctx
        self._rel = crel
        self._root = self  # views share their parent's schema and memory
        self._first = 0    # first row within the root
        self._rows = int(crel.rows)

    @property
    def rows(self):
        return int(self._rel.rows)

    @property
    def types(self):
        return [int(self._rel.cols[i].type) for i in range(self._rel.n_cols)]

    def _c(self):
        return self._rel

    def to_host(self) -> HostRelation:
        n = self.rows
        keys = np.zeros(n, dtype=np.uint64)
        self.ctx._d2h(keys, self._rel.keys)
        cols = []
        for i, t in enumerate(self.types):
            a = np.zeros((n, 2), dtype=np.uint64) if t == U128 else np.zeros(n, dtype=NP_DTYPE[t])
            self.ctx._d2h(a, self._rel.cols[i].data)
            cols.append(a)
        return HostRelation(keys, cols, self.types)

    def slice(self, a, b) -> "DeviceRelation":
        """Rows [a, b) as a view (no ownership; the parent must outlive it)."""
        v = _Relation()
        C.memmove(C.byref(v), C.byref(self._rel), C.sizeof(_Relation))
        v.rows = b - a
        v.keys = self._rel.keys + 8 * a
        for i, t in enumerate(self.types):
            v.cols[i].data = self._rel.cols[i].data + TYPE_SIZE[t] * a
        rel = DeviceRelation(self.ctx, v)
        rel._view = True
        rel._root = self._root
        rel._first = self._first + a
        return rel

    def release(self):
        if getattr(self, "_view", False):
            return
        if self._rel.keys:
            _raise(self.ctx._p, load_library().phj_relation_release(self.ctx._p, C.byref(self._rel)))


class JoinOutput:
    """Join output columns on the host: keys, build-side (left) and probe-side (right)
    columns in schema order (src/JoinOutput.h:18-77)."""

    def __init__(self, keys, left, left_types, right, right_types):
        self.keys, self.left, self.left_types, self.right, self.right_types = keys, left, left_types, right, right_types

    @property
    def rows(self):
        return int(self.keys.shape[0])

    def blocks(self, block_rows=PIPELINE_BLOCK_ROWS):
        """Chunks of at most block_rows rows, like the reference's OutputBlocks."""
        for a in range(0, self.rows, block_rows):
            b = min(self.rows, a + block_rows)
            yield JoinOutput(self.keys[a:b], [c[a:b] for c in self.left], self.left_types,
                             [c[a:b] for c in self.right], self.right_types)


def _batch_descs(pairs, cfgs, tuning):
    """phj_join_desc array for join_batch. When every build side is a view of
    one device relation and every probe side a view of another (the
    back-to-back layout of generate_batch), the descs are filled column-wise
    with numpy from a template desc; otherwise one by one."""
    n = len(pairs)
    descs = (_JoinDesc * max(n, 1))()
    if n == 0:
        return descs
    tmpl = _JoinDesc()
    tmpl.scheme = SCHEME_PHJ_PURE
    tmpl.threads = 1
    for k, v in (tuning or {}).items():
        setattr(tmpl.tuning, k, v)
    b0, p0 = pairs[0]
    tmpl.build = b0._c()
    tmpl.probe = p0._c()
    rb, rp = getattr(b0, "_root", None), getattr(p0, "_root", None)
    views = rb is not None and rp is not None and \
        all(getattr(b, "_root", None) is rb and getattr(p, "_root", None) is rp for b, p in pairs)
    if not views or any(len(c.pass_bits) > MAX_PASSES for c in cfgs):
        for i, (b, p) in enumerate(pairs):
            d = descs[i]
            C.memmove(C.byref(d), C.byref(tmpl), C.sizeof(_JoinDesc))
            d.build = b._c()
            d.probe = p._c()
            cfg = cfgs[i]
            d.n_passes = len(cfg.pass_bits)
            for k, bits in enumerate(cfg.pass_bits):
                d.pass_bits[k] = int(bits)
        return descs
    size = C.sizeof(_JoinDesc)
    buf = np.frombuffer(descs, dtype=np.uint8).reshape(n, size)
    buf[:] = np.frombuffer(C.string_at(C.byref(tmpl), size), dtype=np.uint8)

    def field(off, dtype):
        return np.ndarray((n,), dtype=dtype, buffer=descs, offset=off, strides=(size,))

    for side, root in (("build", rb), ("probe", rp)):
        base = getattr(_JoinDesc, side).offset
        views = [b if side == "build" else p for b, p in pairs]
        first = np.array([v._first for v in views], dtype=np.uint64)
        field(base + _Relation.rows.offset, np.uint64)[:] = [v._rows for v in views]
        r = root._rel
        field(base + _Relation.keys.offset, np.uint64)[:] = np.uint64(r.keys or 0) + np.uint64(8) * first
        for c in range(r.n_cols):
            off = base + _Relation.cols.offset + c * C.sizeof(_Column) + _Column.data.offset
            width = np.uint64(TYPE_SIZE[int(r.cols[c].type)])
            field(off, np.uint64)[:] = np.uint64(r.cols[c].data or 0) + width * first
    field(_JoinDesc.n_passes.offset, np.uint32)[:] = [len(c.pass_bits) for c in cfgs]
    pb = np.zeros((n, MAX_PASSES), dtype=np.uint8)
    for i, c in enumerate(cfgs):
        pb[i, :len(c.pass_bits)] = c.pass_bits
    buf[:, _JoinDesc.pass_bits.offset:_JoinDesc.pass_bits.offset + MAX_PASSES] = pb
    return descs


class JoinResult:
    """ChjResult / PhjResult / PhjBepResult with the output left in HBM."""

    _PHASES = ("build_shuffle", "build", "probe_shuffle", "probe", "eviction_overhead")
    _SCALARS = ("e2e_wall_ms", "peak_mem_bytes", "bep_budget_mib", "bep_evictions", "bep_refinements",
                "bep_peak_bytes", "bep_build_skip_retries")

    def __init__(self, ctx, res: _JoinResult):
        self.ctx = ctx
        self._res = res

    def __getattr__(self, name):
        # phase timings and counters are read from the C struct on first use
        # (a batch returns thousands of results)
        if name in JoinResult._PHASES:
            t = getattr(self._res, name)
            v = PhaseTiming(t.wall_ms, t.ns_per_row)
        elif name in JoinResult._SCALARS:
            v = getattr(self._res, name)
        else:
            raise AttributeError(name)
        self.__dict__[name] = v
        return v

    @property
    def out_rows(self):
        return int(self._res.out_rows)

    def checksum(self):
        cs = _Checksum()
        _raise(self.ctx._p, load_library().phj_result_checksum(self.ctx._p, C.byref(self._res), C.byref(cs)))
        return (int(cs.rows), int(cs.lane0), int(cs.lane1))

    def to_host(self) -> JoinOutput:
        r = self._res
        n = int(r.out_rows)

        def alloc(t):
            return np.zeros((n, 2), dtype=np.uint64) if t == U128 else np.zeros(n, dtype=NP_DTYPE[t])

        lt = [int(r.left_types[i]) for i in range(r.n_left)]
        rt = [int(r.right_types[i]) for i in range(r.n_right)]
        keys = np.zeros(n, dtype=np.uint64)
        left = [alloc(t) for t in lt]
        right = [alloc(t) for t in rt]
        lp = (C.c_void_p * MAX_COLS)(*([a.ctypes.data for a in left] + [None] * (MAX_COLS - len(left))))
        rp = (C.c_void_p * MAX_COLS)(*([a.ctypes.data for a in right] + [None] * (MAX_COLS - len(right))))
        _raise(self.ctx._p, load_library().phj_result_copy_to_host(self.ctx._p, C.byref(r), keys.ctypes.data, lp, rp))
        return JoinOutput(keys, left, lt, right, rt)

    def release(self):
        if self._res.handle:
            _raise(self.ctx._p, load_library().phj_result_release(self.ctx._p, C.byref(self._res)))


def _tuning(d: dict | None) -> _Tuning:
    t = _Tuning()
    for k, v in (d or {}).items():
        setattr(t, k, v)
    return t


class ResidentBuild:
    """A build side partitioned once and kept in HBM (phj_build_register); every
    probe() is runPHJ's probe half against it. Probes may come from any Context
    on the device; release() goes through the registering one."""

    def __init__(self, ctx, ptr, rows):
        self.ctx = ctx
        self._h = ptr
        self.rows = rows

    @property
    def nbytes(self):
        n = C.c_uint64()
        _raise(self.ctx._p, load_library().phj_build_bytes(self._h, C.byref(n)))
        return int(n.value)

    def probe(self, probe, tuning: dict | None = None, ctx=None) -> JoinResult:
        c = ctx or self.ctx
        res = _JoinResult()
        p = probe._c()
        t = _tuning(tuning)
        _raise(c._p, load_library().phj_probe_registered(c._p, self._h, C.byref(p), C.byref(t), C.byref(res)))
        return JoinResult(c, res)

    def release(self):
        if self._h:
            _raise(self.ctx._p, load_library().phj_build_release(self.ctx._p, self._h))
            self._h = None


class Partitioned:
    """Radix shuffle output: contiguous columns grouped by partition; offsets[p]..offsets[p+1]."""

    def __init__(self, ctx, p: _Partitioned):
        self.ctx = ctx
        self._p = p

    @property
    def parts(self):
        return int(self._p.parts)

    @property
    def rows(self):
        return int(self._p.rows)

    def offsets(self):
        a = np.zeros(self.parts + 1, dtype=np.uint64)
        self.ctx._d2h(a, self._p.offsets)
        return a

    def to_host(self):
        n = self.rows
        keys = np.zeros(n, dtype=np.uint64)
        self.ctx._d2h(keys, self._p.keys)
        cols = []
        for i in range(self._p.n_cols):
            t = int(self._p.col_types[i])
            a = np.zeros((n, 2), dtype=np.uint64) if t == U128 else np.zeros(n, dtype=NP_DTYPE[t])
            self.ctx._d2h(a, self._p.cols[i])
            cols.append(a)
        return HostRelation(keys, cols, [int(self._p.col_types[i]) for i in range(self._p.n_cols)]), self.offsets()

    def release(self):
        if self._p.handle:
            _raise(self.ctx._p, load_library().phj_partitioned_release(self.ctx._p, C.byref(self._p)))


def gen_spec(seed=42, build_rows=1 << 20, probe_rows=1 << 24, distribution=DIST_UNIFORM, match_rate=1.0,
             build_types=(U64,), probe_types=(U64,), zipf_max_dup=64, zipf_build_s=1.2, zipf_probe_s=0.8) -> _GenSpec:
    s = _GenSpec()
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


def _mix64(z):
    z = np.uint64(z)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return int(z ^ (z >> np.uint64(31)))


def batched_join_specs(seed=42, n_joins=4096, lo_log2=10, hi_log2=20, probe_factor=8, match_rate=0.75,
                       build_types=(U64,), probe_types=(U64,)):
    """A batch of independent uniform joins: join i has floor(2^U[lo, hi]) distinct
    build keys (U drawn from a counter hash of (seed, i)), probe_factor times as
    many probe rows, match_rate of them matching (misses come from the disjoint
    range [2^63, 2^64)), seeds derived from (seed, i)."""
    specs = []
    for i in range(n_joins):
        u = (_mix64((seed * 0x9E3779B97F4A7C15 + i + 1) & (2**64 - 1)) >> 11) / float(1 << 53)
        rows = int(2.0 ** (lo_log2 + (hi_log2 - lo_log2) * u))
        js = _mix64((seed ^ ((i + 1) * 0xD1B54A32D192ED03)) & (2**64 - 1))
        specs.append(gen_spec(seed=js, build_rows=rows, probe_rows=probe_factor * rows, distribution=DIST_UNIFORM,
                              match_rate=match_rate, build_types=build_types, probe_types=probe_types))
    return specs


def _pass_array(pass_bits):
    pb = list(pass_bits)
    return (C.c_uint8 * max(len(pb), 1))(*[min(max(int(b), 0), 255) for b in pb]), len(pb)


class Context:
    """One device and one HIP stream; owns the device arena."""

    def __init__(self, device=0):
        lib = load_library()
        p = C.c_void_p()
        rc = lib.phj_ctx_create(device, C.byref(p))
        if rc != OK:
            raise PhjError(rc, "no usable HIP device" if rc == ERR_NO_DEVICE else "context creation failed")
        self._p = p

    def close(self):
        if self._p:
            load_library().phj_ctx_destroy(self._p)
            self._p = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- memory
    def _d2h(self, arr: np.ndarray, dev_ptr):
        if arr.nbytes:
            _raise(self._p, load_library().phj_memcpy_d2h(self._p, arr.ctypes.data, dev_ptr, arr.nbytes))

    def mem_stats(self):
        cur, peak = C.c_uint64(), C.c_uint64()
        _raise(self._p, load_library().phj_mem_stats(self._p, C.byref(cur), C.byref(peak)))
        return int(cur.value), int(peak.value)

    def placement_stats(self):
        """(blocks marked slow by the scatter placement watch, bytes marked now)."""
        n, b = C.c_uint64(), C.c_uint64()
        _raise(self._p, load_library().phj_placement_stats(self._p, C.byref(n), C.byref(b)))
        return int(n.value), int(b.value)

    def synchronize(self):
        _raise(self._p, load_library().phj_synchronize(self._p))

    def set_mem_limit(self, nbytes):
        """Cap the device bytes this context's live allocations may hold (0: no cap)."""
        _raise(self._p, load_library().phj_set_mem_limit(self._p, int(nbytes)))

    # -- primitives
    def intHash64Batch(self, keys) -> np.ndarray:
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.empty_like(keys)
        _raise(self._p, load_library().phj_int_hash64_batch(self._p, keys.ctypes.data, keys.size, out.ctypes.data, 0))
        return out

    def radixShuffle(self, rel, cfg: RadixConfig, threads=1, out_schema=None, in_col_index=None) -> Partitioned:
        """radixShuffle (src/RadixPartition.h:187-199). With out_schema and
        in_col_index, output column c is input column in_col_index[c] (whose
        type must be out_schema[c]); without them the columns are kept as they
        are. Each partition holds its rows in input order."""
        out = _Partitioned()
        pb, n = _pass_array(cfg.pass_bits)
        r = rel._c()
        if out_schema is None and in_col_index is None:
            _raise(self._p, load_library().phj_radix_shuffle(self._p, C.byref(r), pb, n, C.byref(out)))
            return Partitioned(self, out)
        out_schema, in_col_index = list(out_schema or []), list(in_col_index or [])
        if len(out_schema) != len(in_col_index):
            raise ValueError("radixShuffle: out_schema and in_col_index size mismatch")  # src/RadixPartition.cpp:559-560
        if any(i < 0 for i in in_col_index):
            raise ValueError("radixShuffle: negative in_col_index")
        types = (C.c_uint32 * max(1, len(out_schema)))(*out_schema)
        index = (C.c_uint32 * max(1, len(in_col_index)))(*in_col_index)
        _raise(self._p, load_library().phj_radix_shuffle_project(self._p, C.byref(r), types, index, len(out_schema), pb,
                                                                 n, C.byref(out)))
        return Partitioned(self, out)

    def refineToLeaves(self, pass1: Partitioned, pass_bits) -> Partitioned:
        out = _Partitioned()
        pb, n = _pass_array(pass_bits)
        _raise(self._p, load_library().phj_refine_to_leaves(self._p, C.byref(pass1._p), pb, n, C.byref(out)))
        return Partitioned(self, out)

    def hashtable_build_find(self, build_keys, probe_keys, block_rows=PIPELINE_BLOCK_ROWS):
        """JoinHashTable insert + batchFind with BlockStore chains: (heads, next) as (n, 2) uint32
        arrays of RowRefCell{block_no, row_no}."""
        bk = np.ascontiguousarray(build_keys, dtype=np.uint64)
        pk = np.ascontiguousarray(probe_keys, dtype=np.uint64)
        heads = np.zeros((pk.size, 2), dtype=np.uint32)
        nxt = np.zeros((bk.size, 2), dtype=np.uint32)
        _raise(self._p, load_library().phj_hashtable_build_find(self._p, bk.ctypes.data, bk.size, pk.ctypes.data, pk.size,
                                                                block_rows, 0, heads.ctypes.data, nxt.ctypes.data))
        return heads, nxt

    def hashtable_build_find_noalloc(self, build_keys, probe_keys, block_rows=PIPELINE_BLOCK_ROWS):
        """hashtable_build_find through the entry point that allocates nothing:
        the keys are copied into device buffers this call allocates up front
        (the caller's inputs and outputs), then the library works in its
        fixed per-GPU scratch. Returns (heads, next) like hashtable_build_find."""
        lib = load_library()
        bk = np.ascontiguousarray(build_keys, dtype=np.uint64)
        pk = np.ascontiguousarray(probe_keys, dtype=np.uint64)
        bufs = []
        try:
            def dev(nbytes):
                ptr = C.c_void_p()
                _raise(self._p, lib.phj_dev_alloc(self._p, max(nbytes, 1), C.byref(ptr)))
                bufs.append(ptr)
                return ptr
            db, dpk, dh, dn = dev(bk.nbytes), dev(pk.nbytes), dev(pk.size * 8), dev(bk.size * 8)
            for d, a in ((db, bk), (dpk, pk)):
                if a.nbytes:
                    _raise(self._p, lib.phj_memcpy_h2d(self._p, d, a.ctypes.data, a.nbytes))
            before = self.mem_stats()
            _raise(self._p, lib.phj_hashtable_build_find_noalloc(self._p, db, bk.size, dpk, pk.size, block_rows, dh, dn))
            if self.mem_stats() != before:
                raise RuntimeError("phj_hashtable_build_find_noalloc changed the context's device memory")
            heads = np.zeros((pk.size, 2), dtype=np.uint32)
            nxt = np.zeros((bk.size, 2), dtype=np.uint32)
            self._d2h(heads, dh)
            self._d2h(nxt, dn)
            return heads, nxt
        finally:
            for b in bufs:
                lib.phj_dev_free(self._p, b)

    # -- schemes
    def join(self, scheme, build, probe, cfg: RadixConfig | None = None, threads=1, bep_budget_mib=512,
             tuning: dict | None = None) -> JoinResult:
        d = _JoinDesc()
        d.scheme = scheme
        d.threads = threads
        d.build = build._c()
        d.probe = probe._c()
        pb = list((cfg or RadixConfig()).pass_bits)
        if len(pb) > MAX_PASSES:
            raise ValueError("at most 8 passes")
        d.n_passes = len(pb)
        for i, b in enumerate(pb):
            d.pass_bits[i] = min(max(int(b), 0), 255)
        d.bep_budget_bytes = int(bep_budget_mib) << 20
        for k, v in (tuning or {}).items():
            setattr(d.tuning, k, v)
        res = _JoinResult()
        _raise(self._p, load_library().phj_join(self._p, C.byref(d), C.byref(res)))
        return JoinResult(self, res)

    def register_build(self, build, cfg: RadixConfig | None = None, tuning: dict | None = None) -> ResidentBuild:
        cfg = cfg or RadixConfig.for_build_rows(build.rows)
        pb, n = _pass_array(cfg.pass_bits)
        r = build._c()
        t = _tuning(tuning)
        h = C.c_void_p()
        _raise(self._p, load_library().phj_build_register(self._p, C.byref(r), pb, n, C.byref(t), C.byref(h)))
        return ResidentBuild(self, h.value, build.rows)

    def join_batch(self, pairs, cfgs=None, tuning: dict | None = None) -> list:
        """Batched PHJ-PURE joins (phj_join_batch): one JoinResult per (build, probe)
        pair, each equal to runPHJ on that pair."""
        n = len(pairs)
        if cfgs is None:
            cfgs = [RadixConfig.for_build_rows(b.rows) for b, _ in pairs]
        descs = _batch_descs(pairs, cfgs, tuning)
        res = (_JoinResult * max(n, 1))()
        _raise(self._p, load_library().phj_join_batch(self._p, n, descs, res))
        # each result is a view of its element of the array (which it keeps alive)
        return [JoinResult(self, res[i]) for i in range(n)]

    def batch_autotune(self, pairs, cfgs=None):
        """phj_batch_autotune: per size class floor(log2(build rows)) present in
        the pairs, the fastest leaf size (build rows) found by timing; kept by
        this context for its later join_batch calls. Returns
        {class: (leaf_rows, best_ms)} for the classes present."""
        n = len(pairs)
        if cfgs is None:
            cfgs = [RadixConfig.for_build_rows(b.rows) for b, _ in pairs]
        descs = _batch_descs(pairs, cfgs, None)
        table = np.zeros(BATCH_CLASSES, dtype=np.uint32)
        ms = np.zeros(BATCH_CLASSES, dtype=np.float64)
        _raise(self._p, load_library().phj_batch_autotune(self._p, n, descs, table.ctypes.data, ms.ctypes.data))
        return {c: (int(table[c]), float(ms[c])) for c in range(BATCH_CLASSES) if ms[c] > 0}

    def batch_set_leaf_rows(self, table):
        """Sets the leaf-size table of join_batch (BATCH_CLASSES entries)."""
        t = np.ascontiguousarray(table, dtype=np.uint32)
        if t.shape != (BATCH_CLASSES,):
            raise ValueError(f"the table has {BATCH_CLASSES} entries")
        _raise(self._p, load_library().phj_batch_set_leaf_rows(self._p, t.ctypes.data))

    def runCHJ(self, build, probe, threads=1, **kw) -> JoinResult:
        return self.join(SCHEME_CHJ, build, probe, None, threads, **kw)

    def runPHJ(self, build, probe, cfg: RadixConfig, threads=1, **kw) -> JoinResult:
        return self.join(SCHEME_PHJ_PURE, build, probe, cfg, threads, **kw)

    def runPhjBep(self, build, probe, cfg: RadixConfig, threads=1, bep_budget_mib=512, **kw) -> JoinResult:
        return self.join(SCHEME_PHJ, build, probe, cfg, threads, bep_budget_mib=bep_budget_mib, **kw)

    # -- workloads
    def generate(self, spec: _GenSpec):
        b, p = _Relation(), _Relation()
        _raise(self._p, load_library().phj_gen_relations(self._p, C.byref(spec), C.byref(b), C.byref(p)))
        return DeviceRelation(self, b), DeviceRelation(self, p)

    def generate_batch(self, specs):
        """Joins generated back to back: (build, probe, build_off, probe_off) with
        join i's rows at [off[i], off[i+1]) of each relation."""
        n = len(specs)
        arr = (_GenSpec * n)(*specs)
        b, p = _Relation(), _Relation()
        boff = np.zeros(n + 1, dtype=np.uint64)
        poff = np.zeros(n + 1, dtype=np.uint64)
        _raise(self._p, load_library().phj_gen_batch(self._p, n, arr, C.byref(b), C.byref(p), boff.ctypes.data,
                                                     poff.ctypes.data))
        return DeviceRelation(self, b), DeviceRelation(self, p), boff, poff

    def expected_checksum(self, spec: _GenSpec):
        cs = _Checksum()
        _raise(self._p, load_library().phj_gen_expected_checksum(self._p, C.byref(spec), C.byref(cs)))
        return (int(cs.rows), int(cs.lane0), int(cs.lane1))

    # -- measurement
    def set_kernel_timing(self, on=True):
        _raise(self._p, load_library().phj_set_kernel_timing(self._p, 1 if on else 0))

    def reset_kernel_times(self):
        _raise(self._p, load_library().phj_kernel_times_reset(self._p))

    def kernel_times(self):
        buf = C.create_string_buffer(4096)
        _raise(self._p, load_library().phj_kernel_names(self._p, buf, 4096))
        out = {}
        for name in filter(None, buf.value.decode().split(",")):
            ms, n = C.c_double(), C.c_uint64()
            _raise(self._p, load_library().phj_kernel_times(self._p, name.encode(), C.byref(ms), C.byref(n)))
            out[name] = (float(ms.value), int(n.value))
        return out

    def stream_bench(self, nbytes, reps=5):
        """Streaming rates of the device, in bytes/s: {"copy": bytes read and
        written per second copying nbytes, "read": reading, "write": writing}."""
        lib = load_library()
        a, b = C.c_void_p(), C.c_void_p()
        _raise(self._p, lib.phj_dev_alloc(self._p, nbytes, C.byref(a)))
        _raise(self._p, lib.phj_dev_alloc(self._p, nbytes, C.byref(b)))
        try:
            out = {}
            for name, mode, moved in (("copy", 0, 2 * nbytes), ("read", 1, nbytes), ("write", 2, nbytes)):
                ms = C.c_double()
                _raise(self._p, lib.phj_stream_bench(self._p, b, a, nbytes, mode, reps, C.byref(ms)))
                out[name] = moved / (ms.value * 1e-3)
            return out
        finally:
            lib.phj_dev_free(self._p, a)
            lib.phj_dev_free(self._p, b)

    PROBE_RANDOM_READ, PROBE_SCRATCH_COPY, PROBE_PASS_COLUMNS, PROBE_PASS_ROWS = 0, 1, 2, 3

    def layout_probe(self, mode, nbytes, param=0, reps=3):
        """Mean ms of one layout-probe pass (phj_layout_probe): PROBE_RANDOM_READ
        (param random 16-byte reads of an nbytes table), PROBE_SCRATCH_COPY
        (nbytes copied through a param-byte scratch), PROBE_PASS_COLUMNS /
        PROBE_PASS_ROWS (one radix pass's pattern over nbytes / 32 rows into
        param partitions)."""
        ms = C.c_double()
        _raise(self._p, load_library().phj_layout_probe(self._p, mode, nbytes, param, reps, C.byref(ms)))
        return float(ms.value)

    PLACEMENT_SEPARATE = (1 << 64) - 1

    def placement_probe(self, rows, sets, reps=3, gap=PLACEMENT_SEPARATE):
        """Mean ms of one radix pass's write pattern (256 partitions, rows of
        8 + 8 + 16 bytes) into each of `sets` freshly allocated sets of output
        columns (phj_placement_probe): each column its own allocation, or one
        allocation per set with `gap` bytes between its columns."""
        ms = (C.c_double * max(sets, 1))()
        _raise(self._p, load_library().phj_placement_probe(self._p, rows, sets, gap, reps, ms))
        return [float(v) for v in ms]

    def copy_bench(self, nbytes, reps=5):
        lib = load_library()
        a, b = C.c_void_p(), C.c_void_p()
        _raise(self._p, lib.phj_dev_alloc(self._p, nbytes, C.byref(a)))
        _raise(self._p, lib.phj_dev_alloc(self._p, nbytes, C.byref(b)))
        try:
            ms = C.c_double()
            _raise(self._p, lib.phj_copy_bench(self._p, b, a, nbytes, reps, C.byref(ms)))
            return float(ms.value)
        finally:
            lib.phj_dev_free(self._p, a)
            lib.phj_dev_free(self._p, b)
