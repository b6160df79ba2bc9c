"""Batched joins (phj_join_batch): every join of a batch must give exactly the
rows of the same join run alone (runPHJ), checked by per-join checksums against
the plain GPU path and, for small joins, against the CPU oracle."""
import numpy as np
import pytest

import phj_gpu as G
import pyoracle as O

pytestmark = pytest.mark.gpu


def pairs_of(build, probe, boff, poff):
    return [(build.slice(int(boff[i]), int(boff[i + 1])), probe.slice(int(poff[i]), int(poff[i + 1])))
            for i in range(len(boff) - 1)]


def plain(ctx, b, p):
    r = ctx.runPHJ(b, p, G.RadixConfig.for_build_rows(b.rows))
    out = (r.out_rows, r.checksum())
    r.release()
    return out


@pytest.mark.parametrize("lo,hi,factor", [(4, 12, 3), (9, 20, 2)])
def test_batch_matches_plain_path(ctx, lo, hi, factor):
    specs = G.batched_join_specs(seed=lo * 100 + hi, n_joins=24, lo_log2=lo, hi_log2=hi, probe_factor=factor,
                                 match_rate=0.75, build_types=(G.U32, G.U64), probe_types=(G.U128,))
    build, probe, boff, poff = ctx.generate_batch(specs)
    pairs = pairs_of(build, probe, boff, poff)
    res = ctx.join_batch(pairs)
    assert len(res) == len(pairs)
    for (b, p), r in zip(pairs, res):
        assert (r.out_rows, r.checksum()) == plain(ctx, b, p)
    # releasing half of a group's results leaves the others intact
    keep = [(r.out_rows, r.checksum()) for r in res[1::2]]
    for r in res[0::2]:
        r.release()
    assert [(r.out_rows, r.checksum()) for r in res[1::2]] == keep
    for r in res[1::2]:
        r.release()
    build.release()
    probe.release()


def test_batch_matches_oracle(ctx):
    specs = G.batched_join_specs(seed=9, n_joins=16, lo_log2=3, hi_log2=13, probe_factor=4, match_rate=0.6,
                                 build_types=(G.U8, G.U64), probe_types=(G.U16,))
    build, probe, boff, poff = ctx.generate_batch(specs)
    res = ctx.join_batch(pairs_of(build, probe, boff, poff))
    for s, r in zip(specs, res):
        ob, op = O.gen(O.make_spec(seed=s.seed, build_rows=s.build_rows, probe_rows=s.probe_rows, distribution=0,
                                   match_rate=s.match_ppm / 1e6, build_types=(G.U8, G.U64), probe_types=(G.U16,)))
        want = O.serialize_sorted(O.reference_join(ob, op))
        assert r.out_rows == want.shape[0]
        assert r.checksum() == O.checksum_rows(want)
        r.release()
    build.release()
    probe.release()


def test_batch_with_repeated_build_keys(ctx):
    """Joins with repeated build keys inside a group are rerun on the plain path."""
    specs = G.batched_join_specs(seed=3, n_joins=6, lo_log2=8, hi_log2=14, probe_factor=4, match_rate=0.8)
    zipf = G.gen_spec(seed=77, build_rows=300, probe_rows=5000, distribution=G.DIST_ZIPF, match_rate=0.9)
    specs = specs[:3] + [zipf] + specs[3:]
    build, probe, boff, poff = ctx.generate_batch(specs)
    pairs = pairs_of(build, probe, boff, poff)
    res = ctx.join_batch(pairs)
    for (b, p), r in zip(pairs, res):
        assert (r.out_rows, r.checksum()) == plain(ctx, b, p)
        r.release()
    build.release()
    probe.release()


def test_batch_of_separate_relations(ctx):
    """Relations that do not lie back to back form groups of one join."""
    rels = [ctx.generate(G.gen_spec(seed=s, build_rows=n, probe_rows=4 * n, match_rate=0.5))
            for s, n in ((1, 100), (2, 5000), (3, 70000))]
    res = ctx.join_batch(rels)
    for (b, p), r in zip(rels, res):
        assert (r.out_rows, r.checksum()) == plain(ctx, b, p)
        r.release()
        b.release()
        p.release()


def test_batch_empty_joins(ctx):
    specs = [G.gen_spec(seed=1, build_rows=0, probe_rows=100), G.gen_spec(seed=2, build_rows=50, probe_rows=0),
             G.gen_spec(seed=3, build_rows=3000, probe_rows=9000, match_rate=0.5)]
    build, probe, boff, poff = ctx.generate_batch(specs)
    pairs = pairs_of(build, probe, boff, poff)
    res = ctx.join_batch(pairs)
    assert [r.out_rows for r in res[:2]] == [0, 0]
    assert (res[2].out_rows, res[2].checksum()) == plain(ctx, *pairs[2])
    for r in res:
        r.release()
    build.release()
    probe.release()


def test_batch_gathers_scattered_small_joins(ctx):
    """Small joins whose relations lie apart (each generated on its own) run
    as one group gathered back to back; a large one among them runs in place;
    every join equals its closed form."""
    specs = [G.gen_spec(seed=40 + i, build_rows=n, probe_rows=5 * n, match_rate=0.6, build_types=(G.U16, G.U64),
                        probe_types=(G.U128,)) for i, n in enumerate((300, 4000, 1, 20000, 2_000_000, 7000, 0, 900))]
    rels = [ctx.generate(s) for s in specs]
    res = ctx.join_batch(rels)
    for s, r in zip(specs, res):
        assert r.checksum() == ctx.expected_checksum(s)
        r.release()
    for b, p in rels:
        b.release()
        p.release()


def test_batch_autotune(ctx):
    """The autotuner picks a leaf size per size class present and the batch
    keeps its results with the tuned table (and with a hand-set one)."""
    specs = G.batched_join_specs(seed=17, n_joins=48, lo_log2=10, hi_log2=17, probe_factor=4, match_rate=0.75)
    build, probe, boff, poff = ctx.generate_batch(specs)
    pairs = pairs_of(build, probe, boff, poff)
    table = ctx.batch_autotune(pairs)
    classes = {int(np.log2(b.rows)) for b, _ in pairs}
    assert set(table) == classes
    assert all(rows in (512, 1024, 2048, 4096) and ms > 0 for rows, ms in table.values())
    for leaf_rows in (None, 600):
        if leaf_rows is not None:
            ctx.batch_set_leaf_rows([leaf_rows] * G.BATCH_CLASSES)
        res = ctx.join_batch(pairs)
        for s, r in zip(specs, res):
            assert r.checksum() == ctx.expected_checksum(s)
            r.release()
    ctx.batch_set_leaf_rows([2048] * G.BATCH_CLASSES)
    build.release()
    probe.release()
