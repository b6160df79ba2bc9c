"""Every benchmark configuration pinned inside the GPU suite: the generators'
closed-form checksums (each matching probe row pairs with the build row it was
drawn from, or with every build row of its distinct key under Zipf) checked
against the CPU oracle at small sizes, then used at the configurations' full
sizes, where the oracle would take too long. Also the boundary cases the
reference accepts: 25..32 radix bits, host probe sides in a batch, and a join
that fails for lack of memory returning everything it held."""
import numpy as np
import pytest

import phj_gpu as G
import pyoracle as O

pytestmark = pytest.mark.gpu


def host_rel(o: O.Relation) -> G.HostRelation:
    return G.HostRelation(o.keys, o.cols, o.types)


ZIPF_SPECS = [dict(seed=1, build_rows=4096, probe_rows=1 << 16, build_types=(0, 1, 3), probe_types=(3,)),
              dict(seed=7, build_rows=300, probe_rows=20000, build_types=(4,), probe_types=(2, 4), zipf_max_dup=5),
              dict(seed=3, build_rows=1, probe_rows=500, build_types=(3,), probe_types=(3,))]


@pytest.mark.parametrize("kw", ZIPF_SPECS)
def test_zipf_closed_form_matches_oracle(ctx, kw):
    """The device's closed form (inverse shuffle by modular inverses) equals the
    oracle's (inverse shuffle tabulated from the forward walk), which equals the
    oracle's hash-map join of the generated rows."""
    dev = ctx.expected_checksum(G.gen_spec(distribution=G.DIST_ZIPF, match_rate=0.9, **kw))
    ospec = O.make_spec(distribution=2, match_rate=0.9, **kw)
    assert dev == O.expected_checksum(ospec)
    b, p = O.gen(ospec)
    assert dev == O.reference_join_checksum(b, p)


def test_zipf_schemes_match_closed_form(ctx):
    spec = G.gen_spec(seed=11, build_rows=3000, probe_rows=200000, distribution=G.DIST_ZIPF, match_rate=0.9,
                      build_types=(G.U8, G.U16, G.U64), probe_types=(G.U64,))
    want = ctx.expected_checksum(spec)
    b, p = ctx.generate(spec)
    for run in (lambda: ctx.runPHJ(b, p, G.RadixConfig([4, 4])), lambda: ctx.runCHJ(b, p),
                lambda: ctx.runPhjBep(b, p, G.RadixConfig([4, 4]), bep_budget_mib=1)):
        r = run()
        assert r.checksum() == want
        r.release()
    b.release()
    p.release()


def test_config2_full_size():
    """2^27 x 2^30, build u32+u64, probe u64+u128, 50% match, pass_bits {8,8}:
    65,536 leaves, a 2^30-row output bound, u32 row offsets near their limit."""
    ctx = G.Context(0)
    spec = G.gen_spec(seed=42, build_rows=1 << 27, probe_rows=1 << 30, match_rate=0.5, build_types=(G.U32, G.U64),
                      probe_types=(G.U64, G.U128))
    b, p = ctx.generate(spec)
    r = ctx.runPHJ(b, p, G.RadixConfig([8, 8]))
    got = (r.out_rows, r.checksum())
    r.release()
    b.release()
    p.release()
    want = ctx.expected_checksum(spec)
    ctx.close()
    assert got == (want[0], want)
    assert abs(want[0] - (1 << 29)) < (1 << 17)


def test_placement_watch_keeps_results():
    """Radix passes over 2^26 rows or more are timed, and a pass much slower than
    the best of its shape steers later pass outputs off its output block. Joins
    repeated in one context stay equal to the closed form whatever the watch
    marks, and what it reports is bounded by the blocks it could have marked."""
    ctx = G.Context(0)
    assert ctx.placement_stats() == (0, 0)
    spec = G.gen_spec(seed=5, build_rows=1 << 24, probe_rows=1 << 26, match_rate=0.5, build_types=(G.U32,),
                      probe_types=(G.U64, G.U128))
    b, p = ctx.generate(spec)
    want = ctx.expected_checksum(spec)
    got = []
    for _ in range(5):
        r = ctx.runPHJ(b, p, G.RadixConfig([8, 8]))
        got.append((r.out_rows, r.checksum()))
        r.release()
    blocks, marked = ctx.placement_stats()
    b.release()
    p.release()
    ctx.close()
    assert got == [(want[0], want)] * 5
    assert 0 <= marked <= blocks * (16 << 26)


def test_config3_full_size():
    """2^22 distinct keys x Zipf(1.2) duplicates on [1,64] (u8+u16+u64) against
    2^27 probes, 90% Zipf(0.8)-ranked (u64): all three schemes, PHJ with a
    32 MiB budget, against the closed form (about 1.2e9 output rows)."""
    ctx = G.Context(0)
    spec = G.gen_spec(seed=42, build_rows=1 << 22, probe_rows=1 << 27, distribution=G.DIST_ZIPF, match_rate=0.9,
                      build_types=(G.U8, G.U16, G.U64), probe_types=(G.U64,), zipf_max_dup=64, zipf_build_s=1.2,
                      zipf_probe_s=0.8)
    want = ctx.expected_checksum(spec)
    assert 1.0e9 < want[0] < 1.5e9
    b, p = ctx.generate(spec)
    cfg = G.RadixConfig([8, 8])
    got = {}
    for name, run in (("phj_pure", lambda: ctx.runPHJ(b, p, cfg)), ("chj", lambda: ctx.runCHJ(b, p)),
                      ("phj", lambda: ctx.runPhjBep(b, p, cfg, bep_budget_mib=32))):
        r = run()
        got[name] = r.checksum()
        if name == "phj":
            assert r.bep_evictions > 0 and r.bep_refinements > 0
        r.release()
    b.release()
    p.release()
    ctx.close()
    assert got == {k: want for k in got}


def test_batch_config4_mix(ctx):
    """Config 4's own mix: |R| = floor(2^U[10,20]) distinct keys, 8x probes, 75%
    match, u64 payload each side; every join against its closed form."""
    specs = G.batched_join_specs(seed=42, n_joins=1024)
    build, probe, boff, poff = ctx.generate_batch(specs)
    pairs = [(build.slice(int(boff[i]), int(boff[i + 1])), probe.slice(int(poff[i]), int(poff[i + 1])))
             for i in range(len(specs))]
    assert min(b.rows for b, _ in pairs) >= 1 << 10 and max(b.rows for b, _ in pairs) >= 1 << 19
    res = ctx.join_batch(pairs)
    for s, r in zip(specs, res):
        assert r.checksum() == ctx.expected_checksum(s)
        r.release()
    build.release()
    probe.release()


@pytest.mark.parametrize("bits", [[16, 16], [13, 12], [9, 8, 8], [16, 9]])
def test_radix_configs_beyond_device_bits(ctx, bits):
    """25..32 total bits are valid in the reference (src/CLI.cpp:350-355); PHJ
    and PHJ-PURE join through the device's 21-bit plan with the same rows."""
    spec = O.make_spec(seed=19, build_rows=20000, probe_rows=90000, distribution=0, match_rate=0.6,
                       build_types=(G.U32,), probe_types=(G.U64, G.U8))
    ob, op = O.gen(spec)
    want = O.serialize_sorted(O.reference_join(ob, op))
    for r in (ctx.runPHJ(host_rel(ob), host_rel(op), G.RadixConfig(bits)),
              ctx.runPhjBep(host_rel(ob), host_rel(op), G.RadixConfig(bits), bep_budget_mib=1)):
        out = r.to_host()
        got = O.serialize_sorted(O.JoinRows(out.keys, out.left, out.left_types, out.right, out.right_types))
        assert np.array_equal(got, want)
        r.release()


def test_radix_shuffle_beyond_device_bits_unsupported(ctx):
    rel = G.HostRelation(np.arange(100, dtype=np.uint64))
    with pytest.raises(G.PhjError) as e:
        ctx.radixShuffle(rel, G.RadixConfig([16, 16]))
    assert e.value.code == G.ERR_UNSUPPORTED


def test_batch_with_host_probe_sides(ctx):
    """Device build sides with host probe sides: the joins run on the plain
    path (which stages the probe side), each equal to the oracle."""
    specs = [O.make_spec(seed=s, build_rows=n, probe_rows=3 * n, distribution=0, match_rate=0.7, build_types=(G.U64,),
                         probe_types=(G.U32,)) for s, n in ((1, 500), (2, 3000), (3, 9000))]
    pairs, wants, devs = [], [], []
    for s in specs:
        ob, op = O.gen(s)
        gb = G.gen_spec(seed=s.seed, build_rows=s.build_rows, probe_rows=0, build_types=(G.U64,), probe_types=())
        db, dp = ctx.generate(gb)
        dp.release()
        devs.append(db)
        pairs.append((db, host_rel(op)))
        wants.append(O.checksum_rows(O.serialize_sorted(O.reference_join(ob, op))))
    res = ctx.join_batch(pairs)
    for r, w in zip(res, wants):
        assert r.checksum() == w
        r.release()
    for d in devs:
        d.release()


def test_failed_join_returns_its_memory():
    """A join that runs out of memory part way (a context memory limit that
    admits the partition buffers but not the output) raises MemoryError and
    leaves the context holding what it held before."""
    ctx = G.Context(0)
    spec = G.gen_spec(seed=5, build_rows=1 << 20, probe_rows=1 << 23, match_rate=0.5, build_types=(G.U64,),
                      probe_types=(G.U64, G.U128))
    b, p = ctx.generate(spec)
    base, _ = ctx.mem_stats()
    # 300 MiB beyond the inputs: PHJ-PURE's partition buffers (16 + 256 MiB) and
    # CHJ's tables (~44 MiB) fit, but not with the 2^23-row output bound (40 B/row)
    ctx.set_mem_limit(base + (300 << 20))
    for run in (lambda: ctx.runPHJ(b, p, G.RadixConfig([8])), lambda: ctx.runCHJ(b, p),
                lambda: ctx.runPhjBep(b, p, G.RadixConfig([8]), bep_budget_mib=1 << 14)):
        with pytest.raises(MemoryError):
            run()
        assert ctx.mem_stats()[0] == base
    ctx.set_mem_limit(0)
    r = ctx.runPHJ(b, p, G.RadixConfig([8]))
    assert r.checksum() == ctx.expected_checksum(spec)
    r.release()
    assert ctx.mem_stats()[0] == base
    b.release()
    p.release()
    ctx.close()
