"""GPU parity tests: every hot-path primitive and scheme through the C ABI,
checked against the CPU oracle (oracle/) on the same seeded inputs."""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import phj_gpu as G
import pyoracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def host_rel(o: O.Relation) -> G.HostRelation:
    return G.HostRelation(o.keys, o.cols, o.types)


def orc_rel(h: G.HostRelation) -> O.Relation:
    return O.Relation(h.keys, h.cols, h.types)


def sorted_rows_of(out: G.JoinOutput):
    return O.serialize_sorted(O.JoinRows(out.keys, out.left, out.left_types, out.right, out.right_types))


def random_relation(rng, n, types, key_pool=None):
    keys = rng.integers(0, 2**64, n, dtype=np.uint64) if key_pool is None else rng.choice(key_pool, n)
    cols = []
    for t in types:
        if t == G.U128:
            cols.append(rng.integers(0, 2**64, (n, 2), dtype=np.uint64))
        else:
            cols.append(rng.integers(0, np.iinfo(G.NP_DTYPE[t]).max, n, dtype=G.NP_DTYPE[t], endpoint=True))
    return O.Relation(keys, cols, list(types))


# ---------------------------------------------------------------- a1/a2 hash
@pytest.mark.parametrize("n", [0, 1, 3, 1000, 1 << 20])
def test_int_hash64_batch(ctx, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 2**64, n, dtype=np.uint64)
    if n >= 3:
        x[:3] = [0, 2**64 - 1, 42]
    got = ctx.intHash64Batch(x)
    assert np.array_equal(got, O.int_hash64(x))


# ---------------------------------------------------------------- generators
@pytest.mark.parametrize("dist,kw", [
    (G.DIST_UNIFORM, dict(match_rate=0.5, build_types=(2, 3), probe_types=(3, 4))),
    (G.DIST_REFERENCE, dict(build_types=(0, 1, 2), probe_types=(4,))),
    (G.DIST_ZIPF, dict(match_rate=0.9, build_types=(0, 1, 3), probe_types=(3,))),
])
def test_generator_matches_oracle(ctx, dist, kw):
    args = dict(seed=7, build_rows=3000, probe_rows=20000, **kw)
    b, p = ctx.generate(G.gen_spec(distribution=dist, **args))
    ob, op = O.gen(O.make_spec(distribution=dist, **args))
    for dev, ref in ((b, ob), (p, op)):
        h = dev.to_host()
        assert h.rows == ref.rows
        assert np.array_equal(h.keys, ref.keys)
        for a, c in zip(h.cols, ref.cols):
            assert np.array_equal(a, c)
    b.release()
    p.release()


# ---------------------------------------------------------------- a3-a7 partition
def leaf_ids(keys, total_bits):
    return (O.int_hash64(keys) >> np.uint64(64 - total_bits)).astype(np.int64)


def rows_multiset(rel: G.HostRelation, idx=None):
    o = orc_rel(rel) if idx is None else orc_rel(rel).take(idx)
    j = O.JoinRows(o.keys, o.cols, o.types, [], [])
    return O.serialize_sorted(j)


@pytest.mark.parametrize("pass_bits", [[8], [4, 4], [8, 8], [3, 3, 3], [5], [16], [1], [12, 7]])
@pytest.mark.parametrize("n", [0, 1, 4097, 100_000])
def test_radix_shuffle_partitions(ctx, pass_bits, n):
    rng = np.random.default_rng(len(pass_bits) * 1000 + n)
    rel = random_relation(rng, n, [G.U8, G.U32, G.U128])
    part = ctx.radixShuffle(host_rel(rel), G.RadixConfig(pass_bits))
    out, off = part.to_host()
    total = sum(pass_bits)
    assert part.parts == 1 << total
    assert off[0] == 0 and off[-1] == n and np.all(np.diff(off.astype(np.int64)) >= 0)
    # every row sits in the partition of its top `total` hash bits
    lid = leaf_ids(out.keys, total)
    expect = np.repeat(np.arange(part.parts), np.diff(off.astype(np.int64)))
    assert np.array_equal(lid, expect)
    # histogram equals the brute-force top-bit count
    counts = np.bincount(leaf_ids(rel.keys, total), minlength=part.parts) if n else np.zeros(part.parts)
    assert np.array_equal(np.diff(off.astype(np.int64)), counts)
    # the multiset of rows is preserved
    assert np.array_equal(rows_multiset(out), rows_multiset(host_rel(rel)))
    # and every partition holds its rows in input order, row for row as the
    # oracle (pinned to the compiled reference's radixShuffle) places them
    want, woff = O.radix_shuffle(rel, pass_bits)
    assert np.array_equal(off.astype(np.int64), woff)
    assert np.array_equal(out.keys, want.keys)
    for a, b in zip(out.cols, want.cols):
        assert np.array_equal(a, b)
    part.release()


@pytest.mark.parametrize("index", [[], [2], [3, 0], [1, 1, 2], [2, 3, 1, 0]])
@pytest.mark.parametrize("pass_bits", [[8], [5, 5]])
def test_radix_shuffle_projection(ctx, index, pass_bits):
    """radixShuffle(input, out_schema, in_col_index, cfg) (src/RadixPartition.h:187-193):
    columns dropped, reordered and repeated, row for row as the oracle."""
    rng = np.random.default_rng(len(index) * 10 + len(pass_bits))
    rel = random_relation(rng, 30_001, [G.U8, G.U64, G.U128, G.U32])
    types = [rel.types[i] for i in index]
    part = ctx.radixShuffle(host_rel(rel), G.RadixConfig(pass_bits), out_schema=types, in_col_index=index)
    out, off = part.to_host()
    want, woff = O.radix_shuffle(rel, pass_bits, types, index)
    assert out.types == types
    assert np.array_equal(off.astype(np.int64), woff) and np.array_equal(out.keys, want.keys)
    for a, b in zip(out.cols, want.cols):
        assert np.array_equal(a, b)
    part.release()


def test_radix_shuffle_projection_rejects_bad_maps(ctx):
    rel = host_rel(random_relation(np.random.default_rng(2), 10, [G.U64, G.U16]))
    for types, index in (([G.U64], [2]), ([G.U32], [0]), ([G.U64, G.U16], [0]), ([G.U64] * 9, [0] * 9)):
        with pytest.raises(ValueError):
            ctx.radixShuffle(rel, G.RadixConfig([8]), out_schema=types, in_col_index=index)


@pytest.mark.parametrize("pass_bits", [[9], [8, 8], [1]])
def test_radix_shuffle_stable_under_hot_keys(ctx, pass_bits):
    """Most rows share a few keys (one digit takes most of every tile): rows
    keep their input order inside each partition."""
    rng = np.random.default_rng(99)
    n = 200_003
    keys = np.where(rng.random(n) < 0.9, rng.choice(np.array([7, 2**64 - 1, 0], dtype=np.uint64), n),
                    rng.integers(0, 2**64, n, dtype=np.uint64)).astype(np.uint64)
    rel = O.Relation(keys, [np.arange(n, dtype=np.uint32)], [G.U32])
    part = ctx.radixShuffle(host_rel(rel), G.RadixConfig(pass_bits))
    out, off = part.to_host()
    want, woff = O.radix_shuffle(rel, pass_bits)
    assert np.array_equal(off.astype(np.int64), woff)
    assert np.array_equal(out.cols[0], want.cols[0])  # the row numbers: input order in every partition
    part.release()


def test_radix_shuffle_rejects_bad_config(ctx):
    rel = host_rel(random_relation(np.random.default_rng(1), 10, [G.U64]))
    for bad in ([], [0], [17], [16, 16, 1]):
        with pytest.raises(ValueError):
            ctx.radixShuffle(rel, G.RadixConfig(bad))


@pytest.mark.parametrize("pass_bits", [[4, 4], [6, 5], [3, 3, 3], [8], [8, 8]])
def test_refine_to_leaves(ctx, pass_bits):
    """Pass-1 buckets refined to leaves (src/RadixPartition.cpp:417-539): every
    row sits in the leaf of its top hash bits as the oracle's intHash64 gives
    them, and each leaf holds exactly the input rows with that leaf id."""
    rng = np.random.default_rng(sum(pass_bits))
    rel = random_relation(rng, 50_000, [G.U16, G.U64, G.U128])
    total = sum(pass_bits)
    p1 = ctx.radixShuffle(host_rel(rel), G.RadixConfig(pass_bits[:1]))
    leaves = ctx.refineToLeaves(p1, pass_bits)
    assert leaves.parts == 1 << total
    out, off = leaves.to_host()
    lens = np.diff(off.astype(np.int64))
    assert np.array_equal(leaf_ids(out.keys, total), np.repeat(np.arange(leaves.parts), lens))
    want_lid = leaf_ids(rel.keys, total)
    assert np.array_equal(lens, np.bincount(want_lid, minlength=leaves.parts))
    # per-leaf row multisets: rows ordered by (oracle leaf id, row bytes) on both sides
    order = np.argsort(want_lid, kind="stable")
    want_rows = [rows_multiset(host_rel(rel), order[int(off[L]):int(off[L + 1])])
                 for L in range(0, leaves.parts, max(1, leaves.parts // 97))]
    got_rows = [rows_multiset(out, np.arange(int(off[L]), int(off[L + 1])))
                for L in range(0, leaves.parts, max(1, leaves.parts // 97))]
    for a, b in zip(got_rows, want_rows):
        assert np.array_equal(a, b)
    # and the whole relation's multiset
    assert np.array_equal(rows_multiset(out), rows_multiset(host_rel(rel)))
    # stable passes: the leaves hold the rows exactly as one radixShuffle over all passes
    want, woff = O.radix_shuffle(rel, pass_bits)
    assert np.array_equal(off.astype(np.int64), woff) and np.array_equal(out.keys, want.keys)
    for a, b in zip(out.cols, want.cols):
        assert np.array_equal(a, b)
    for x in (p1, leaves):
        x.release()


# ---------------------------------------------------------------- a9-a12 hash table
def test_hashtable_chains(ctx):
    rng = np.random.default_rng(5)
    pool = rng.integers(0, 2**64, 3000, dtype=np.uint64)
    pool[:2] = [2**64 - 1, 0]  # the empty-slot sentinel and zero are ordinary keys
    bkeys = rng.choice(pool, 20_000)
    pkeys = np.concatenate([pool, rng.integers(0, 2**64, 1000, dtype=np.uint64)])
    block_rows = 10_000
    heads, nxt = ctx.hashtable_build_find(bkeys, pkeys, block_rows)

    def row(ref):
        return None if ref[0] == G.INVALID_BLOCK else int(ref[0]) * block_rows + int(ref[1])

    where = {}
    for i, k in enumerate(bkeys.tolist()):
        where.setdefault(k, set()).add(i)
    for j, k in enumerate(pkeys.tolist()):
        chain = set()
        r = row(heads[j])
        steps = 0
        while r is not None:
            assert r not in chain
            chain.add(r)
            r = row(nxt[r])
            steps += 1
            assert steps <= len(bkeys)
        assert chain == where.get(k, set()), j


# ---------------------------------------------------------------- schemes
SCHEMES = [("chj", None), ("phj_pure", [8]), ("phj_pure", [8, 8]), ("phj_pure", [4, 4]), ("phj_pure", [2]),
           ("phj", [8, 8]), ("phj", [6])]


def run_scheme(ctx, name, cfg, b, p, **kw):
    if name == "chj":
        return ctx.runCHJ(b, p, **kw)
    if name == "phj_pure":
        return ctx.runPHJ(b, p, G.RadixConfig(cfg), **kw)
    return ctx.runPhjBep(b, p, G.RadixConfig(cfg), bep_budget_mib=1, **kw)


def datasets():
    rng = np.random.default_rng(11)
    out = {}
    s = O.make_spec(seed=3, build_rows=5000, probe_rows=40000, distribution=0, match_rate=0.5, build_types=(2, 3),
                    probe_types=(3, 4))
    out["uniform_half"] = O.gen(s)
    s = O.make_spec(seed=4, build_rows=3000, probe_rows=30000, distribution=1, build_types=(0, 4), probe_types=(1, 2))
    out["reference_streams"] = O.gen(s)
    s = O.make_spec(seed=5, build_rows=500, probe_rows=30000, distribution=2, match_rate=0.9, build_types=(0, 1, 3),
                    probe_types=(3,))
    out["zipf_multi"] = O.gen(s)
    pool = rng.integers(0, 2**64, 200, dtype=np.uint64)
    pool[:3] = [2**64 - 1, 0, 1]
    out["sentinel_dups"] = (random_relation(rng, 2000, [G.U64], pool), random_relation(rng, 5000, [G.U8], pool))
    out["empty_build"] = (random_relation(rng, 0, [G.U32]), random_relation(rng, 1000, [G.U64]))
    out["empty_probe"] = (random_relation(rng, 1000, [G.U32]), random_relation(rng, 0, [G.U64]))
    out["all_miss"] = (random_relation(rng, 1000, [G.U64]), random_relation(rng, 3000, [G.U64]))
    out["no_payload"] = (random_relation(rng, 1000, [], pool), random_relation(rng, 3000, [], pool))
    out["wide_payload"] = (random_relation(rng, 3000, [0, 1, 2, 3, 4, 4, 3, 0], pool),
                           random_relation(rng, 4000, [4, 3, 2, 1, 0, 4, 4, 4], pool))
    return out


DATA = None


def data():
    global DATA
    if DATA is None:
        DATA = datasets()
    return DATA


@pytest.mark.parametrize("dsname", ["uniform_half", "reference_streams", "zipf_multi", "sentinel_dups", "empty_build",
                                    "empty_probe", "all_miss", "no_payload", "wide_payload"])
@pytest.mark.parametrize("scheme,cfg", SCHEMES)
def test_join_matches_reference(ctx, dsname, scheme, cfg):
    ob, op = data()[dsname]
    want = O.serialize_sorted(O.reference_join(ob, op))
    res = run_scheme(ctx, scheme, cfg, host_rel(ob), host_rel(op))
    out = res.to_host()
    assert out.rows == want.shape[0]
    assert np.array_equal(sorted_rows_of(out), want)
    assert res.checksum() == O.checksum_rows(want)
    res.release()


@pytest.mark.parametrize("flags", [dict(flags=1), dict(lds_leaf_rows=8), dict(ht_load_pct=80),
                                   dict(join_threads=256), dict(join_threads=1024), dict(probe_chunk_rows=1000),
                                   dict(flags=4), dict(flags=8), dict(flags=12), dict(flags=0x10),
                                   dict(flags=0x12), dict(flags=0x40), dict(flags=0x50)])
def test_phj_pure_tuning_paths(ctx, flags):
    ob, op = data()["zipf_multi"]
    want = O.serialize_sorted(O.reference_join(ob, op))
    res = ctx.runPHJ(host_rel(ob), host_rel(op), G.RadixConfig([4, 4]), tuning=flags)
    assert np.array_equal(sorted_rows_of(res.to_host()), want)
    res.release()


def unique_with_sentinels(rng, n_build, n_probe):
    """Distinct build keys including 0 and 2^64-1, probe keys half hits."""
    keys = np.unique(rng.integers(0, 2**64, n_build, dtype=np.uint64, endpoint=False))
    keys = np.unique(np.concatenate([keys, np.array([0, 2**64 - 1], dtype=np.uint64)]))
    rng.shuffle(keys)
    b = random_relation(rng, keys.size, [G.U32, G.U128])
    b = O.Relation(keys, b.cols, b.types)
    hit = rng.random(n_probe) < 0.5
    pk = np.where(hit, keys[rng.integers(0, keys.size, n_probe)], rng.integers(0, 2**64, n_probe, dtype=np.uint64))
    pk[:4] = [2**64 - 1, 0, 2**64 - 1, 5]
    p = random_relation(rng, n_probe, [G.U16, G.U64])
    return b, O.Relation(pk.astype(np.uint64), p.cols, p.types)


@pytest.mark.parametrize("pass_bits", [[4], [2, 2], [8]])
@pytest.mark.parametrize("flags", [0, 2])
def test_fused_and_two_phase_leaf_join(ctx, pass_bits, flags):
    """Unique build keys (the fused leaf join) against the forced two-phase
    join, with leaves whose probe side spans several work items and the
    empty-key sentinel on both sides."""
    rng = np.random.default_rng(21)
    ob, op = unique_with_sentinels(rng, 4000, 300000)
    want = O.serialize_sorted(O.reference_join(ob, op))
    res = ctx.runPHJ(host_rel(ob), host_rel(op), G.RadixConfig(pass_bits), tuning=dict(flags=flags))
    assert res.out_rows == want.shape[0]
    assert np.array_equal(sorted_rows_of(res.to_host()), want)
    res.release()
    res = ctx.runPhjBep(host_rel(ob), host_rel(op), G.RadixConfig(pass_bits), bep_budget_mib=1, tuning=dict(flags=flags))
    assert res.checksum() == O.checksum_rows(want)
    # 1 MiB buffers ~29K probe rows: the 300K probe rows are refined and
    # evicted many times over
    assert res.bep_refinements > (1 << pass_bits[0]) and res.bep_evictions > (1 << sum(pass_bits))
    res.release()


@pytest.mark.parametrize("case", ["unique", "repeated_keys", "oversize_leaves"])
@pytest.mark.parametrize("pass_bits", [[4], [2, 2], [8, 8]])
def test_narrow_probe_side(ctx, case, pass_bits):
    """PHJ_TUNE_NARROW_PROBE: the probe side partitioned as (key, row index),
    probe columns gathered from the unpartitioned relation. Unique build keys
    (with the 0 and 2^64-1 sentinels) take the narrow path; repeated build keys
    and leaves beyond the LDS limit fall back to the full-width join. Rows equal
    the oracle's in every case."""
    rng = np.random.default_rng(37)
    if case == "repeated_keys":
        ob, op = data()["zipf_multi"]
    else:
        ob, op = unique_with_sentinels(rng, 5000, 120000)
    tuning = dict(flags=0x20)
    if case == "oversize_leaves":
        tuning["lds_leaf_rows"] = 8
    want = O.serialize_sorted(O.reference_join(ob, op))
    res = ctx.runPHJ(host_rel(ob), host_rel(op), G.RadixConfig(pass_bits), tuning=tuning)
    assert res.out_rows == want.shape[0]
    assert np.array_equal(sorted_rows_of(res.to_host()), want)
    res.release()


@pytest.mark.parametrize("half_empty", [False, True])
def test_bep_sparse_rounds_read_tables_in_hbm(ctx, half_empty):
    """PHJ best-effort under a 1 MiB budget against 256 leaves of ~1,200 build
    rows: an eviction round holds far fewer probe rows per leaf than build rows,
    so the leaf tables are built once up front (the build phase) and every
    round looks up in them where they lie in HBM. With half_empty, no build key
    hashes into the upper half of the leaves, so half the probe rows (and
    possibly the sentinels) land in leaves without a table."""
    rng = np.random.default_rng(23)
    ob, op = unique_with_sentinels(rng, 600000 if half_empty else 300000, 600000)
    if half_empty:
        keep = (O.int_hash64(ob.keys) >> np.uint64(63)) == 0
        ob = O.Relation(ob.keys[keep], [c[keep] for c in ob.cols], ob.types)
    want = O.checksum_rows(O.serialize_sorted(O.reference_join(ob, op)))
    res = ctx.runPhjBep(host_rel(ob), host_rel(op), G.RadixConfig([4, 4]), bep_budget_mib=1)
    assert res.build.wall_ms > 0  # tables built once
    assert res.checksum() == want
    # 600K probe rows of 18 bytes through 512 KiB of unrefined buffer and
    # 640 KiB of leaf buffer: every bucket and leaf is refined and evicted
    # many times
    assert res.bep_refinements > 16 * 5 and res.bep_evictions > 256 * 5
    assert 0 < res.bep_peak_bytes <= (1 << 20)
    res.release()


def bep_groups(keys, bits):
    return np.unique(O.int_hash64(keys) >> np.uint64(64 - bits)).size


@pytest.mark.parametrize("pass_bits", [[4, 4], [8, 8], [6], [3, 3, 3]])
def test_bep_schedule_counters(ctx, pass_bits):
    """The best-effort schedule (src/PHJBep.cpp:397-695) in its counters: with a
    budget that buffers every probe row, each non-empty pass-1 bucket is
    refined once and each non-empty leaf evicted once (the final refine and
    drain); with a small budget they are refined and evicted repeatedly, the
    buffered leaf rows stay within the budget, and a smaller budget never
    evicts less. The rows equal the oracle's every time."""
    spec = O.make_spec(seed=41, build_rows=40000, probe_rows=400000, distribution=0, match_rate=0.7,
                       build_types=(G.U32,), probe_types=(G.U64, G.U16))
    ob, op = O.gen(spec)
    want = O.checksum_rows(O.serialize_sorted(O.reference_join(ob, op)))
    total = sum(pass_bits)
    buckets, leaves = bep_groups(op.keys, pass_bits[0]), bep_groups(op.keys, total)
    rb = 8 + 8 + 2
    seen = []
    for mib in (4096, 4, 1):
        r = ctx.runPhjBep(host_rel(ob), host_rel(op), G.RadixConfig(pass_bits), bep_budget_mib=mib)
        assert r.checksum() == want
        assert r.bep_budget_mib == mib and r.bep_build_skip_retries == 0
        if mib == 4096:
            assert (r.bep_refinements, r.bep_evictions, r.bep_peak_bytes) == (buckets, leaves, 0)
        else:
            assert r.bep_refinements >= buckets and r.bep_evictions > leaves
            # leaf rows are sampled after each batch, below the leaf high-water
            # mark plus one refine's rows (at most the unrefined high-water mark)
            assert 0 < r.bep_peak_bytes <= (mib << 20) * 5 // 8 + (mib << 20) // 2 + rb * 4096
            assert r.eviction_overhead.wall_ms > 0
        seen.append(r.bep_evictions)
        r.release()
    assert seen == sorted(seen)


@pytest.mark.parametrize("scheme,cfg", [("chj", None), ("phj_pure", [8, 8]), ("phj_pure", [4]), ("phj", [8, 8])])
def test_one_key_with_many_build_rows(ctx, scheme, cfg):
    """A key with 20,000 build rows (its leaf cannot fit LDS however many
    partition bits are used: the HBM-table path with the key's rows laid out as
    one group) beside ordinary keys, a second repeated key and the empty-key
    sentinel, against the oracle row for row."""
    rng = np.random.default_rng(77)
    hot, warm = np.uint64(0x1234_5678_9ABC_DEF0), np.uint64(2**64 - 1)
    bkeys = rng.integers(0, 2**64, 30_000, dtype=np.uint64)
    bkeys[:20_000] = hot
    bkeys[20_000:20_300] = warm
    rng.shuffle(bkeys)
    plain = bkeys[(bkeys != hot) & (bkeys != warm)]
    pkeys = np.concatenate([np.full(60, hot), np.full(40, warm), rng.choice(plain, 3000),
                            rng.integers(0, 2**64, 2000, dtype=np.uint64)])
    rng.shuffle(pkeys)
    ob = O.Relation(bkeys, [rng.integers(0, 2**32, bkeys.size, dtype=np.uint32),
                            rng.integers(0, 2**64, (bkeys.size, 2), dtype=np.uint64)], [G.U32, G.U128])
    op = O.Relation(pkeys, [rng.integers(0, 2**16, pkeys.size, dtype=np.uint16)], [G.U16])
    want = O.serialize_sorted(O.reference_join(ob, op))
    res = run_scheme(ctx, scheme, cfg, host_rel(ob), host_rel(op))
    out = res.to_host()
    assert out.rows == want.shape[0] and want.shape[0] > 1_200_000
    assert np.array_equal(sorted_rows_of(out), want)
    assert res.checksum() == O.checksum_rows(want)
    res.release()


@pytest.mark.parametrize("dsname", ["zipf_multi", "sentinel_dups", "no_payload", "wide_payload"])
@pytest.mark.parametrize("pass_bits", [[4, 4], [6], [8, 8]])
def test_bep_group_images_many_rounds(ctx, dsname, pass_bits):
    """PHJ best-effort against group images (repeated build keys: every key's
    rows adjacent, cells {key, start, count}) under a budget small enough that
    the probe rows are refined and evicted in many rounds, each round joined
    against the images built once. Keys include 0 and the all-ones sentinel
    (sentinel_dups); rows equal the oracle's, sorted byte for byte."""
    ob, op = data()[dsname]
    # the probe side repeated to 120K rows: 1 MiB of buffers holds a fraction of it
    reps = -(-120_000 // op.keys.size)
    tile = lambda a: np.tile(a, (reps,) + (1,) * (a.ndim - 1))[:120_000]
    op = O.Relation(tile(op.keys), [tile(c) for c in op.cols], op.types)
    want = O.serialize_sorted(O.reference_join(ob, op))
    res = ctx.runPhjBep(host_rel(ob), host_rel(op), G.RadixConfig(pass_bits), bep_budget_mib=1)
    assert res.out_rows == want.shape[0]
    assert np.array_equal(sorted_rows_of(res.to_host()), want)
    assert res.build.wall_ms > 0  # the images were built up front
    assert res.bep_evictions >= bep_groups(op.keys, sum(pass_bits))
    res.release()


@pytest.mark.parametrize("chunk", [700, 5000])
def test_chains_across_work_items(ctx, chunk):
    """Repeated build keys in leaves of thousands of rows (tables built by many
    wavefronts, so chain order varies between builds) whose probe side spans
    many work items: each item must expand chains from its own table."""
    spec = O.make_spec(seed=12, build_rows=6000, probe_rows=120000, distribution=2, match_rate=0.9,
                       build_types=(G.U16, G.U64), probe_types=(G.U32,))
    ob, op = O.gen(spec)
    want = O.serialize_sorted(O.reference_join(ob, op))
    for bits in ([4], [2, 3]):  # leaves of about 3,800 and 1,900 build rows, in LDS (an explicit LDS limit
        # keeps the configuration as given)
        res = ctx.runPHJ(host_rel(ob), host_rel(op), G.RadixConfig(bits),
                         tuning=dict(probe_chunk_rows=chunk, lds_leaf_rows=4096))
        assert res.out_rows == want.shape[0]
        assert np.array_equal(sorted_rows_of(res.to_host()), want)
        res.release()


@pytest.mark.parametrize("bits", [[1], [2], [3]])
def test_long_chains_in_emit_batches(ctx, bits):
    """300 keys repeated 40 times each, every probe row matching: a round of
    2,048 probe rows yields ~80,000 multi-match output rows, far more than one
    LDS batch of the cooperative chain emission, and chains of 40 rows run
    across batch boundaries. With an LDS leaf limit of 2,048 rows (which
    keeps the configuration as given), leaves of ~6,000 and ~3,000 build rows
    ([1], [2]) go to HBM tables and ~1,500 ([3]) stay in LDS."""
    rng = np.random.default_rng(77)
    distinct = rng.permutation(np.unique(rng.integers(1, 2**63, 400, dtype=np.uint64))[:300])
    bkeys = rng.permutation(np.repeat(distinct, 40))
    bcol = np.arange(bkeys.size, dtype=np.uint32)
    pkeys = distinct[rng.integers(0, distinct.size, 20000)]
    pcol = rng.integers(0, 2**63, (pkeys.size, 2), dtype=np.uint64)
    ob = O.Relation(bkeys, [bcol], [G.U32])
    op = O.Relation(pkeys, [pcol], [G.U128])
    want = O.serialize_sorted(O.reference_join(ob, op))
    assert want.shape[0] == 20000 * 40
    res = ctx.runPHJ(host_rel(ob), host_rel(op), G.RadixConfig(bits), tuning=dict(lds_leaf_rows=2048))
    assert res.out_rows == want.shape[0]
    assert np.array_equal(sorted_rows_of(res.to_host()), want)
    res.release()


@pytest.mark.parametrize("bits", [[1], [2, 2], [8, 8]])
def test_coarse_radix_config_refined(ctx, bits):
    """A configuration with leaves too large for LDS joins through LDS-sized
    leaves (the output does not depend on them): same rows as the oracle and
    as the configuration kept exactly (HBM leaf tables)."""
    spec = O.make_spec(seed=31, build_rows=60000, probe_rows=400000, distribution=0, match_rate=0.6,
                       build_types=(G.U32, G.U64), probe_types=(G.U64, G.U128))
    ob, op = O.gen(spec)
    want = O.serialize_sorted(O.reference_join(ob, op))
    refined = ctx.runPHJ(host_rel(ob), host_rel(op), G.RadixConfig(bits))
    exact = ctx.runPHJ(host_rel(ob), host_rel(op), G.RadixConfig(bits), tuning=dict(lds_leaf_rows=60000))
    assert np.array_equal(sorted_rows_of(refined.to_host()), want)
    assert refined.checksum() == exact.checksum() == O.checksum_rows(want)
    refined.release()
    exact.release()


def test_join_rejects_bad_radix(ctx):
    ob, op = data()["uniform_half"]
    with pytest.raises(ValueError):
        ctx.runPHJ(host_rel(ob), host_rel(op), G.RadixConfig([20, 20]))
    with pytest.raises(ValueError):
        ctx.runPhjBep(host_rel(ob), host_rel(op), G.RadixConfig([]))


# ---------------------------------------------------------------- checksums
def test_expected_checksum_closed_form(ctx):
    args = dict(seed=9, build_rows=1 << 16, probe_rows=1 << 19, match_rate=0.5, build_types=(2, 3), probe_types=(3, 4))
    spec = G.gen_spec(distribution=G.DIST_UNIFORM, **args)
    gpu_expected = ctx.expected_checksum(spec)
    cpu_expected = O.expected_checksum(O.make_spec(distribution=0, **args), threads=8)
    assert gpu_expected == cpu_expected
    b, p = ctx.generate(spec)
    for scheme, cfg in (("phj_pure", [8, 8]), ("chj", None), ("phj", [8, 8])):
        res = run_scheme(ctx, scheme, cfg, b, p)
        assert res.checksum() == cpu_expected, scheme
        res.release()
    b.release()
    p.release()


# ---------------------------------------------------------------- C++ adapter (drop-in signatures)
def _write_side(f, rel: O.Relation):
    f.write(struct.pack("<QI", rel.rows, len(rel.types)))
    for t in rel.types:
        f.write(struct.pack("<I", t))
    f.write(rel.keys.tobytes())
    for c in rel.cols:
        f.write(np.ascontiguousarray(c).tobytes())


@pytest.mark.parametrize("scheme,pass_bits,callers", [(1, [8], 1), (3, [8], 1), (3, [4, 4], 1), (2, [4, 4], 1),
                                                      (3, [8, 8], 2), (1, [8], 3), (2, [6], 2)])
def test_cpp_adapter_drop_in(scheme, pass_bits, callers):
    """The drop-in signatures through the C++ adapter, from `callers` threads at
    once (the reference's scheme functions may run concurrently): every
    caller's rows equal the oracle's."""
    exe = os.path.join(ROOT, "phj-bench_amd", "host", "adapter_selftest")
    ob, op = data()["reference_streams"]
    want = O.serialize_sorted(O.reference_join(ob, op))
    with tempfile.TemporaryDirectory() as td:
        inp, outp = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(inp, "wb") as f:
            pb = bytes(pass_bits + [0] * (8 - len(pass_bits)))
            f.write(struct.pack("<QII", 0x31464C45534A4850, scheme, len(pass_bits)) + pb + struct.pack("<QQ", 1, 4))
            _write_side(f, ob)
            _write_side(f, op)
        r = subprocess.run([exe, inp, outp, str(callers)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        with open(outp, "rb") as f:
            n, rs = struct.unpack("<QQ", f.read(16))
            got = np.frombuffer(f.read(), dtype=np.uint8).reshape(n, rs)
    assert np.array_equal(got, want)


def test_stream_bench_rates(ctx):
    """The streaming-rate harness runs each mode and reports plausible rates
    (above 100 GB/s, below the 8 TB/s peak by any margin of timing noise)."""
    rates = ctx.stream_bench(256 << 20, reps=3)
    assert set(rates) == {"copy", "read", "write"}
    for v in rates.values():
        assert 1e11 < v < 1e13


def test_layout_probe_modes(ctx):
    """The layout probes run each mode with plausible timings and refuse
    arguments outside their documented ranges."""
    # 2^20 random reads of a 1 MiB table, of a 64 MiB table
    for mib in (1, 64):
        ms = ctx.layout_probe(G.Context.PROBE_RANDOM_READ, mib << 20, 1 << 20, reps=2)
        assert 0 < ms < 1000
    # 64 MiB through a 4 MiB scratch, and through one that holds it all
    for scratch in (4 << 20, 64 << 20):
        ms = ctx.layout_probe(G.Context.PROBE_SCRATCH_COPY, 64 << 20, scratch, reps=2)
        assert 0 < ms < 1000
    # one pass over 2^20 rows of 32 bytes: 32 MiB read and written at > 100 GB/s
    for mode in (G.Context.PROBE_PASS_COLUMNS, G.Context.PROBE_PASS_ROWS):
        for parts in (1, 256, 4096):
            ms = ctx.layout_probe(mode, 32 << 20, parts, reps=2)
            assert 0 < ms < 2 * (32 << 20) / 1e11 * 1e3
    for mode, nbytes, param in ((G.Context.PROBE_RANDOM_READ, 1024, 0), (G.Context.PROBE_SCRATCH_COPY, 1 << 20, 0),
                                (G.Context.PROBE_PASS_COLUMNS, 32 << 20, 3), (G.Context.PROBE_PASS_ROWS, 32 << 20, 8192),
                                (G.Context.PROBE_PASS_ROWS, 1024, 256), (7, 1 << 20, 1)):
        with pytest.raises(ValueError):
            ctx.layout_probe(mode, nbytes, param, reps=1)


def test_placement_probe(ctx):
    """The placement probe times the pass pattern into each freshly allocated
    set and refuses out-of-range arguments."""
    for gap in (ctx.PLACEMENT_SEPARATE, 0, 4096):
        ms = ctx.placement_probe(1 << 20, 2, reps=2, gap=gap)
        assert len(ms) == 2 and all(0 < v < 2 * (32 << 20) / 1e11 * 1e3 for v in ms)
    for rows, sets in ((1000, 1), (1 << 20, 0), (1 << 20, 9)):
        with pytest.raises(ValueError):
            ctx.placement_probe(rows, sets, reps=1)


@pytest.mark.parametrize("build_types", [[G.U128], [G.U16, G.U32], [G.U8, G.U64], [G.U32, G.U16, G.U8],
                                         [G.U64, G.U64], [G.U8, G.U128], [G.U64, G.U128, G.U64]])
def test_chj_inline_cell_layouts(ctx, build_types):
    """CHJ with the build columns carried inline in 32- and 64-byte cells (and
    16-byte cells when they do not fit): every layout of the inline bytes, one
    row per key and a few repeated keys, against the oracle."""
    rng = np.random.default_rng(len(build_types) * 7 + build_types[0])
    keys = np.unique(rng.integers(0, 2**64, 6000, dtype=np.uint64))
    keys = np.concatenate([keys, keys[:50]])  # 50 keys with two build rows
    rng.shuffle(keys)
    b = random_relation(rng, keys.size, build_types)
    b = O.Relation(keys, b.cols, b.types)
    pk = np.where(rng.random(20000) < 0.7, rng.choice(keys, 20000), rng.integers(0, 2**64, 20000, dtype=np.uint64))
    p = random_relation(rng, pk.size, [G.U64])
    p = O.Relation(pk, p.cols, p.types)
    want = O.serialize_sorted(O.reference_join(b, p))
    res = ctx.runCHJ(host_rel(b), host_rel(p))
    assert np.array_equal(sorted_rows_of(res.to_host()), want)
    res.release()
