"""Resident build sides and the serving loop: every served request's result
equals the closed form of its generator spec (the exact join output), alone
and in the mix, and a resident probe equals runPHJ on the same pair."""
import threading
import time

import numpy as np
import pytest

import phj_gpu as G
import serve

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bits", [[8], [4, 4], [3]])
def test_resident_probe_matches_runphj(ctx, bits):
    spec = G.gen_spec(seed=5, build_rows=20000, probe_rows=150000, match_rate=0.7, build_types=(G.U32, G.U128),
                      probe_types=(G.U64,))
    b, p = ctx.generate(spec)
    res = ctx.register_build(b, G.RadixConfig(bits))
    assert res.nbytes >= 20000 * 28
    want = ctx.runPHJ(b, p, G.RadixConfig(bits))
    for _ in range(2):  # the resident side survives repeated probes
        got = res.probe(p)
        assert (got.out_rows, got.checksum()) == (want.out_rows, want.checksum())
        assert got.build_shuffle.wall_ms == 0.0
        got.release()
    assert want.checksum() == ctx.expected_checksum(spec)
    want.release()
    # a probe from another context on the same device
    other = G.Context(0)
    got = res.probe(p, ctx=other)
    assert got.checksum() == ctx.expected_checksum(spec)
    got.release()
    other.close()
    res.release()
    b.release()
    p.release()


def test_resident_probe_with_repeated_keys(ctx):
    spec = G.gen_spec(seed=8, build_rows=400, probe_rows=20000, distribution=G.DIST_ZIPF, match_rate=0.9)
    b, p = ctx.generate(spec)
    res = ctx.register_build(b, G.RadixConfig([2, 2]))
    want = ctx.runPHJ(b, p, G.RadixConfig([2, 2]))
    got = res.probe(p)
    assert (got.out_rows, got.checksum()) == (want.out_rows, want.checksum())
    for r in (want, got):
        r.release()
    res.release()
    b.release()
    p.release()


def test_serving_loop_results_alone_and_in_mix(ctx):
    kinds, residents, _ = serve.request_pool(ctx, seed=3, n_new=4, n_resident=3, n_resident_probes=3, lo_log2=8,
                                             hi_log2=15, probe_lo=2, probe_hi=6)
    exact = [serve.expected(ctx, k) for k in kinds]
    alone = [serve.run_request(ctx, k, want_checksum=True) for k in kinds]
    assert alone == exact
    srv = serve.JoinServer(workers=3, budget_bytes=1 << 30)
    sched = serve.poisson_schedule(rate=2000, n_requests=60, n_kinds=len(kinds), seed=4)
    outs = srv.run(kinds, sched, verify_every=1)
    assert {o.kind for o in outs} == set(range(len(kinds)))
    for o in outs:
        assert (o.out_rows, o.checksum) == exact[o.kind]
        assert o.done >= o.admitted >= o.arrival - 1e-6
    summary = serve.summarize(outs, kinds)
    assert summary["requests"] == 60 and summary["p99_ms"] >= summary["p50_ms"] > 0
    # a tiny budget admits one request at a time and still completes all
    srv_small = serve.JoinServer(workers=2, budget_bytes=1)
    outs = srv_small.run(kinds, sched[:10], verify_every=1)
    assert all((o.out_rows, o.checksum) == exact[o.kind] for o in outs)
    for s in (srv, srv_small):
        s.close()
    for k in kinds:
        for rel in (k.build, k.probe):
            if rel is not None:
                rel.release()
    for r in residents:
        r.release()


def test_resident_release_refused_while_probing(ctx):
    """Release of a resident build side fails while a probe of it runs on
    another thread (counted on the host), and succeeds once it has returned.
    The probe (2^29 rows) runs for tens of milliseconds; the release is tried
    a few milliseconds after the probing thread has entered the call."""
    spec = G.gen_spec(seed=21, build_rows=1 << 20, probe_rows=1 << 29, match_rate=0.5)
    b, p = ctx.generate(spec)
    res = ctx.register_build(b)
    other = G.Context(0)
    out, entering = {}, threading.Event()

    def prober():
        entering.set()
        r = res.probe(p, ctx=other)
        out["t"] = time.perf_counter()
        out["cs"] = r.checksum()
        r.release()

    th = threading.Thread(target=prober)
    th.start()
    entering.wait(timeout=60)
    time.sleep(0.003)
    t_try = time.perf_counter()
    with pytest.raises(ValueError, match="in flight"):
        res.release()
    th.join()
    assert out["t"] > t_try  # the probe was still running when release was refused
    assert out["cs"] == ctx.expected_checksum(spec)
    res.release()
    assert res._h is None
    other.close()
    b.release()
    p.release()


def test_serving_batches_queued_small_requests(ctx):
    """A burst of arrivals queues many small requests: workers admit same-schema
    small requests together and join them in one phj_join_batch call; every
    served result still equals its closed form, batched or not."""
    kinds, residents, n_new = serve.request_pool(ctx, seed=13, n_new=5, n_resident=2, n_resident_probes=2, lo_log2=9,
                                                 hi_log2=14, probe_lo=2, probe_hi=6)
    exact = [serve.expected(ctx, k) for k in kinds]
    srv = serve.JoinServer(workers=2, budget_bytes=1 << 32)
    sched = serve.poisson_schedule(rate=200000, n_requests=120, n_kinds=len(kinds), seed=5, n_new=n_new,
                                   resident_share=0.2)
    outs = srv.run(kinds, sched, verify_every=1)
    assert srv.batches > 0 and srv.batched_requests >= 2 * srv.batches
    for o in outs:
        assert (o.out_rows, o.checksum) == exact[o.kind]
    srv.close()
    for k in kinds:
        for rel in (k.build, k.probe):
            if rel is not None:
                rel.release()
    for r in residents:
        r.release()
