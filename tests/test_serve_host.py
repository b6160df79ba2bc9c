"""Host logic of the serving loop and the batch workload (no GPU)."""
import numpy as np

import phj_gpu as G
import serve


def test_poisson_schedule_rate_and_resident_share():
    sched = serve.poisson_schedule(rate=1000, n_requests=20000, n_kinds=48, seed=1, n_new=32, resident_share=0.3)
    t = np.array([a for a, _ in sched])
    k = np.array([b for _, b in sched])
    assert np.all(np.diff(t) > 0)
    assert abs(len(t) / t[-1] - 1000) < 30  # mean inter-arrival 1 ms
    share = np.mean(k >= 32)
    assert abs(share - 0.3) < 0.02
    assert k.min() >= 0 and k.max() < 48


def test_summarize_latency_percentiles():
    kinds = [serve.RequestInput(None, None, None, G.RadixConfig([1]), 10, 90, 0)]
    outs = []
    for i in range(100):
        o = serve.Outcome(i, 0, arrival=i * 0.01)
        o.done = o.arrival + (i + 1) * 1e-3
        outs.append(o)
    sm = serve.summarize(outs, kinds)
    assert sm["requests"] == 100
    assert abs(sm["p50_ms"] - 50.5) < 0.6
    assert sm["p99_ms"] > 98
    assert abs(sm["tuples_per_s"] - 100 * 100 / sm["makespan_s"]) < 1e-6


def test_batched_join_specs_size_law():
    specs = G.batched_join_specs(seed=42, n_joins=4096)
    r = np.array([s.build_rows for s in specs])
    assert r.min() >= 1024 and r.max() <= 1 << 20
    # floor(2^U[10,20]) has mean (2^20 - 2^10) / (10 ln 2)
    assert abs(r.mean() / ((2**20 - 2**10) / (10 * np.log(2))) - 1) < 0.05
    assert all(s.probe_rows == 8 * s.build_rows and s.match_ppm == 750000 for s in specs)
    assert len({s.seed for s in specs}) == 4096


def test_radix_config_for_build_rows():
    for rows in (0, 1, 2048, 2049, 5000, 1 << 19, (1 << 20) + 5, 1 << 27):
        cfg = G.RadixConfig.for_build_rows(rows)
        total = cfg.total_bits()
        assert all(1 <= b <= 8 for b in cfg.pass_bits)
        assert total >= 1 and (rows >> total) <= 2048 or total == 16
        if total > 1:
            assert (rows >> (total - 1)) > 2048
