"""The allocation-free JoinHashTable entry point
(phj_hashtable_build_find_noalloc) against the allocating one
(phj_hashtable_build_find) and brute force, on two calls in a row with the
second smaller, so scratch contents left by the first call cannot leak into
the second."""
import numpy as np
import pytest

import phj_gpu as G

pytestmark = pytest.mark.gpu


def chains(heads, nxt, block_rows):
    """Per probe row, the set of build rows its chain enumerates."""
    def row(ref):
        return None if ref[0] == G.INVALID_BLOCK else int(ref[0]) * block_rows + int(ref[1])

    out = []
    for h in heads:
        seen, r = set(), row(h)
        while r is not None:
            assert r not in seen and len(seen) <= nxt.shape[0]
            seen.add(r)
            r = row(nxt[r])
        out.append(seen)
    return out


def keys_case(rng, n_build, n_distinct, n_probe):
    pool = rng.integers(0, 2**64, n_distinct, dtype=np.uint64)
    pool[:2] = [2**64 - 1, 0]  # the empty-slot sentinel and zero are ordinary keys
    bkeys = rng.choice(pool, n_build)
    pkeys = np.concatenate([pool, rng.integers(0, 2**64, n_probe - n_distinct, dtype=np.uint64)])
    return bkeys, pkeys


def test_noalloc_matches_allocating_entry_point():
    ctx = G.Context(0)
    rng = np.random.default_rng(29)
    block_rows = 10_000
    for n_build, n_distinct, n_probe in ((300_000, 120_000, 1_500_000), (20_000, 3_000, 9_000)):
        bkeys, pkeys = keys_case(rng, n_build, n_distinct, n_probe)
        heads, nxt = ctx.hashtable_build_find_noalloc(bkeys, pkeys, block_rows)
        want_h, want_n = ctx.hashtable_build_find(bkeys, pkeys, block_rows)
        got = chains(heads, nxt, block_rows)
        assert got == chains(want_h, want_n, block_rows)
        where = {}
        for i, k in enumerate(bkeys.tolist()):
            where.setdefault(k, set()).add(i)
        assert got == [where.get(k, set()) for k in pkeys.tolist()]
        # keys that occur once have exactly one possible head
        once = np.array([len(c) == 1 for c in got])
        assert np.array_equal(heads[once], want_h[once])
    ctx.close()


def test_noalloc_rejects_oversize_build():
    ctx = G.Context(0)
    with pytest.raises(G.PhjError) as e:
        ctx.hashtable_build_find_noalloc(np.arange((1 << 20) + 1, dtype=np.uint64), np.arange(4, dtype=np.uint64))
    assert e.value.code == G.ERR_CAPACITY
    # the largest accepted build side still works
    bkeys = np.arange(1 << 20, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    heads, _ = ctx.hashtable_build_find_noalloc(bkeys, bkeys[:1000])
    assert np.array_equal(heads[:, 0].astype(np.int64) * 10_000 + heads[:, 1], np.arange(1000))
    ctx.close()
