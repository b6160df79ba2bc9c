"""CPU tests pinning the oracle: closed-form known answers, brute force on tiny
inputs, the committed golden fixtures (written by the compiled reference), and
the compiled reference itself (oracle/_ref) when it is built here."""
import json
import os

import numpy as np
import pytest

import pyoracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="compiled reference oracle/_ref not built")


# ---------------------------------------------------------------- intHash64 (src/Hash.h:18-26)
KATS = {0: 0x28BFAA0AFBB6AA0B, 1: 0x70C030F292395DD8, 2: 0xAA41A1291D9862CC, 2**64 - 1: 0x36137209C4D8F428,
        42: 0x5427C49F446106AA}


def test_int_hash64_known_answers():
    for x, want in KATS.items():
        assert O.lib().orc_int_hash64(x) == want
        assert O.int_hash64_py(x) == want
    assert np.array_equal(O.int_hash64(np.array(list(KATS), dtype=np.uint64)),
                          np.array(list(KATS.values()), dtype=np.uint64))


def test_int_hash64_matches_python_closed_form():
    rng = np.random.default_rng(0)
    x = rng.integers(0, 2**64, 2000, dtype=np.uint64)
    assert [int(v) for v in O.int_hash64(x)] == [O.int_hash64_py(int(v)) for v in x]


def test_int_hash64_is_a_bijection_sample():
    x = np.arange(1 << 20, dtype=np.uint64)
    assert np.unique(O.int_hash64(x)).size == x.size


def test_int_hash64_golden():
    with open(os.path.join(GOLDEN, "hash_kat.json")) as f:
        g = json.load(f)
    x = np.array([int(v, 16) for v in g["in"]], dtype=np.uint64)
    want = np.array([int(v, 16) for v in g["out"]], dtype=np.uint64)
    assert np.array_equal(O.int_hash64(x), want)


@needs_ref
def test_int_hash64_matches_compiled_reference_avx512():
    rng = np.random.default_rng(3)
    x = rng.integers(0, 2**64, 100_003, dtype=np.uint64)
    assert np.array_equal(O.int_hash64(x), O.ref_int_hash64(x))


# ---------------------------------------------------------------- reference join (src/Reference.cpp)
def brute_force_join(b: O.Relation, p: O.Relation):
    pi, bi = [], []
    for j, k in enumerate(p.keys.tolist()):
        for r, bk in enumerate(b.keys.tolist()):
            if bk == k:
                pi.append(j)
                bi.append(r)
    return np.array(pi, dtype=np.int64), np.array(bi, dtype=np.int64)


def small_case(seed, nb, np_, pool_size, btypes, ptypes):
    rng = np.random.default_rng(seed)
    pool = rng.integers(0, 2**64, pool_size, dtype=np.uint64)
    pool[0] = 2**64 - 1

    def rel(n, types):
        cols = [rng.integers(0, 2**64, (n, 2), dtype=np.uint64) if t == 4 else
                rng.integers(0, 256 ** O.TYPE_SIZE[t] - 1, n, dtype=O.NP_DTYPE[t], endpoint=True) for t in types]
        return O.Relation(rng.choice(pool, n), cols, types)

    return rel(nb, btypes), rel(np_, ptypes)


@pytest.mark.parametrize("seed", range(5))
def test_reference_join_matches_brute_force(seed):
    b, p = small_case(seed, 60, 150, 40, [2, 4], [0, 3])
    pi, bi = O.reference_join_pairs(b, p)
    bpi, bbi = brute_force_join(b, p)
    # same multiset of (probe, build) pairs; the reference emits in probe order
    assert sorted(zip(pi.tolist(), bi.tolist())) == sorted(zip(bpi.tolist(), bbi.tolist()))
    assert np.all(np.diff(pi.astype(np.int64)) >= 0)
    # within one probe row, the build rows come in insertion order (vector push_back)
    for j in np.unique(pi):
        rows = bi[pi == j]
        assert np.all(np.diff(rows.astype(np.int64)) > 0)


def test_serialisation_layout():
    b = O.Relation(np.array([5], dtype=np.uint64), [np.array([0x01020304], dtype=np.uint32)], [2])
    p = O.Relation(np.array([5], dtype=np.uint64), [np.array([[7, 9]], dtype=np.uint64)], [4])
    rows = O.serialize_sorted(O.reference_join(b, p))
    expect = (5).to_bytes(8, "little") + (0x01020304).to_bytes(4, "little") + (7).to_bytes(8, "little") + \
        (9).to_bytes(8, "little")
    assert rows.tobytes() == expect


def test_checksum_definition_and_invariance():
    b, p = small_case(9, 300, 900, 100, [0, 1, 4], [2, 3])
    j = O.reference_join(b, p)
    rows = O.serialize_sorted(j)
    cs = O.checksum(j)
    assert cs == O.checksum_rows(rows)
    assert cs == O.reference_join_checksum(b, p)
    # matches the pure-Python restatement row by row
    a = c = 0
    for r in rows[:200]:
        la, lc = O.row_digest_py(r.tobytes())
        a, c = (a + la) % 2**64, (c + lc) % 2**64
    part = O.checksum_rows(rows[:200])
    assert part == (200, a, c)
    # order-independent, but sensitive to one flipped byte
    perm = np.random.default_rng(1).permutation(rows.shape[0])
    assert O.checksum_rows(rows[perm]) == cs
    bad = rows.copy()
    bad[3, 5] ^= 1
    assert O.checksum_rows(bad) != cs


# ---------------------------------------------------------------- generators
def test_uniform_generator_invariants():
    s = O.make_spec(seed=42, build_rows=1 << 16, probe_rows=1 << 18, distribution=0, match_rate=0.5,
                    build_types=(2, 3), probe_types=(3, 4))
    b, p = O.gen(s)
    assert np.unique(b.keys).size == b.rows
    assert b.keys.min() >= 1 and b.keys.max() < 2**63
    hit = p.keys < np.uint64(2**63)
    assert np.isin(p.keys[hit], b.keys).all()
    assert 0.49 < hit.mean() < 0.51
    # closed-form checksum equals the joined checksum
    assert O.expected_checksum(s) == O.reference_join_checksum(b, p)


def test_zipf_generator_invariants():
    s = O.make_spec(seed=1, build_rows=4096, probe_rows=1 << 16, distribution=2, match_rate=0.9,
                    build_types=(0, 1, 3), probe_types=(3,))
    b, p = O.gen(s)
    keys, counts = np.unique(b.keys, return_counts=True)
    assert keys.size == 4096 and counts.min() >= 1 and counts.max() <= 64
    assert 8.5 < counts.mean() < 11.5  # E[k] = 10.03 for Zipf(1.2) on [1, 64]
    hit = p.keys < np.uint64(2**63)
    assert 0.88 < hit.mean() < 0.92


@pytest.mark.parametrize("seed,distinct,probes,maxdup", [(1, 4096, 1 << 16, 64), (7, 300, 20000, 5), (3, 1, 500, 64),
                                                          (9, 5000, 0, 64)])
def test_zipf_closed_form_checksum(seed, distinct, probes, maxdup):
    """The Zipf closed form (each hit pairs with every build row of its
    distinct key, found through the inverse of the build-row shuffle) equals
    the checksum of the hash-map join over the generated rows."""
    s = O.make_spec(seed=seed, build_rows=distinct, probe_rows=probes, distribution=2, match_rate=0.9,
                    build_types=(0, 1, 3), probe_types=(3, 4), zipf_max_dup=maxdup)
    b, p = O.gen(s)
    want = O.reference_join_checksum(b, p)
    assert O.expected_checksum(s) == want
    # a join that drops one build row of a repeated key must not pass
    keys, counts = np.unique(b.keys, return_counts=True)
    if probes and counts.max() > 1:
        drop = int(np.flatnonzero(b.keys == keys[np.argmax(counts)])[0])
        keep = np.arange(b.rows) != drop
        short = O.Relation(b.keys[keep], [c[keep] for c in b.cols], b.types)
        assert O.reference_join_checksum(short, p) != want


def test_golden_reference_join():
    """Fixture written by the compiled reference (tests/golden/make_golden.py)."""
    g = np.load(os.path.join(GOLDEN, "ref_join_small.npz"))
    b = O.Relation(g["b_keys"], [g["b_c0"], g["b_c1"]], [int(t) for t in g["b_types"]])
    p = O.Relation(g["p_keys"], [g["p_c0"], g["p_c1"]], [int(t) for t in g["p_types"]])
    rows = O.serialize_sorted(O.reference_join(b, p))
    assert np.array_equal(rows, g["rows"])
    assert list(O.checksum_rows(rows)) == [int(v) for v in g["checksum"]]


def test_golden_reference_streams():
    """The reference DataGen streams restated by the oracle equal the reference's own."""
    g = np.load(os.path.join(GOLDEN, "ref_datagen.npz"))
    s = O.make_spec(seed=int(g["seed"]), build_rows=int(g["b_keys"].size), probe_rows=int(g["p_keys"].size),
                    distribution=1, build_types=tuple(int(t) for t in g["b_types"]),
                    probe_types=tuple(int(t) for t in g["p_types"]))
    b, p = O.gen(s)
    assert np.array_equal(b.keys, g["b_keys"]) and np.array_equal(p.keys, g["p_keys"])
    for i, c in enumerate(b.cols):
        assert np.array_equal(c, g[f"b_c{i}"])
    for i, c in enumerate(p.cols):
        assert np.array_equal(c, g[f"p_c{i}"])


def test_golden_schemes_multi_match():
    """All three reference schemes agree on a multi-match fixture with sentinel keys."""
    g = np.load(os.path.join(GOLDEN, "ref_multi.npz"))
    b = O.Relation(g["b_keys"], [g["b_c0"]], [int(g["b_types"][0])])
    p = O.Relation(g["p_keys"], [g["p_c0"]], [int(g["p_types"][0])])
    cs = O.reference_join_checksum(b, p)
    for name in ("reference", "chj", "phj", "phj_pure"):
        assert [int(v) for v in g[f"cs_{name}"]] == list(cs), name


@needs_ref
@pytest.mark.parametrize("dist", [0, 1, 2])
def test_oracle_matches_compiled_reference(dist):
    s = O.make_spec(seed=5, build_rows=2000, probe_rows=20000, distribution=dist, match_rate=0.7,
                    build_types=(2, 4), probe_types=(0, 3))
    b, p = O.gen(s)
    want = O.ref_reference_checksum(b, p)
    assert O.reference_join_checksum(b, p) == want
    rows = O.serialize_sorted(O.reference_join(b, p))
    assert np.array_equal(rows, O.ref_reference_rows(b, p, rows.shape[1], rows.shape[0]))
    for scheme, pb in ((1, (8,)), (3, (4, 4)), (2, (4, 4))):
        assert O.ref_run(scheme, b, p, pb, threads=2, bep_budget_mib=1).checksum.tuple() == want


# ---------------------------------------------------------------- radixShuffle (src/RadixPartition.cpp:546-577)
def stable_passes_py(keys, pass_bits):
    """Pass-by-pass bucketing in pure Python on a small input: each pass splits
    every bucket of the previous pass by the next hash bits, appending rows in
    order. Returns the row order and the bucket sizes."""
    buckets = [list(range(len(keys)))]
    consumed = 0
    for bits in pass_bits:
        shift = 64 - consumed - bits
        nxt = []
        for b in buckets:
            sub = [[] for _ in range(1 << bits)]
            for i in b:
                sub[(O.int_hash64_py(int(keys[i])) >> shift) & ((1 << bits) - 1)].append(i)
            nxt.extend(sub)
        buckets = nxt
        consumed += bits
    return [i for b in buckets for i in b], [len(b) for b in buckets]


@pytest.mark.parametrize("pass_bits", [[1], [3], [2, 3], [2, 2, 2], [5, 1]])
def test_radix_shuffle_matches_pass_by_pass_bucketing(pass_bits):
    rel, _ = small_case(sum(pass_bits), 700, 1, 50, [1, 4], [3])
    got, off = O.radix_shuffle(rel, pass_bits)
    order, sizes = stable_passes_py(rel.keys, pass_bits)
    assert np.array_equal(got.keys, rel.keys[order])
    assert np.array_equal(got.cols[1], rel.cols[1][order])
    assert np.array_equal(np.diff(off), sizes)


def rel_for_shuffle(seed, n):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 2**64, n, dtype=np.uint64)
    keys[: n // 4] = keys[n // 2: n // 2 + n // 4]  # repeated keys stay in input order
    types = [0, 3, 4, 2]
    cols = [rng.integers(0, 2**64, (n, 2), dtype=np.uint64) if t == 4 else
            rng.integers(0, 256 ** O.TYPE_SIZE[t] - 1, n, dtype=O.NP_DTYPE[t], endpoint=True) for t in types]
    return O.Relation(keys, cols, types)


@needs_ref
@pytest.mark.parametrize("n", [0, 1, 9_999, 10_001, 45_000])
@pytest.mark.parametrize("pass_bits", [[8], [4, 4], [3, 3, 3], [7, 6]])
@pytest.mark.parametrize("threads", [1, 3])
def test_radix_shuffle_matches_compiled_reference(n, pass_bits, threads):
    """Row for row, with the reference's 10,000-row blocks spread over threads."""
    rel = rel_for_shuffle(n + len(pass_bits), n)
    want, woff = O.ref_radix_shuffle(rel, pass_bits, threads=threads)
    got, off = O.radix_shuffle(rel, pass_bits)
    assert np.array_equal(off, woff)
    assert np.array_equal(got.keys, want.keys)
    for a, b in zip(got.cols, want.cols):
        assert np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("index", [[], [2], [3, 0], [1, 1, 2], [2, 3, 1, 0]])
def test_radix_shuffle_projection_matches_compiled_reference(index):
    rel = rel_for_shuffle(11, 23_456)
    types = [rel.types[i] for i in index]
    want, woff = O.ref_radix_shuffle(rel, [5, 5], types, index, threads=2)
    got, off = O.radix_shuffle(rel, [5, 5], types, index)
    assert np.array_equal(off, woff) and np.array_equal(got.keys, want.keys)
    assert got.types == types == want.types
    for a, b in zip(got.cols, want.cols):
        assert np.array_equal(a, b)


@needs_ref
def test_reference_radix_shuffle_rejects_bad_configs():
    rel = rel_for_shuffle(1, 100)
    for bad in ([], [0]):  # src/RadixPartition.cpp:552-556
        with pytest.raises(ValueError):
            O.ref_radix_shuffle(rel, bad)
