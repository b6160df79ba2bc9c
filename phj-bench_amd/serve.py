"""Serving loop over the join library (the caller of runPHJ in a server).

Requests arrive open-loop (Poisson); each is a PHJ-PURE join whose build side
is either new (partitioned by the request) or one of the device-resident build
sides registered up front (phj_build_register), whose probe half alone runs
(phj_probe_registered). Admission is by HBM footprint: a request holds a
reservation (its output bound and partition buffers) from admission until it
completes; arrivals wait in FIFO order while the budget is full. W workers,
one HIP stream (Context) each, admit and run requests, so up to W joins are in
flight on the GPU at once; a worker admits the next request as soon as its
current one finishes.

Small requests with their own build sides are batched continuously: a worker
admitting one also admits the queued requests of the same schema that fit
the budget and joins them in one phj_join_batch call (JoinServer).

Latency is completion minus arrival. The reference has no serving mode; its
unit of work, runPHJ (src/PHJ.h:31), is what every request runs.
"""
import threading
import time
from dataclasses import dataclass

import numpy as np

import phj_gpu as G

_PAYLOAD_TYPES = (G.U32, G.U64, G.U128)


@dataclass
class RequestInput:
    """Inputs of one request kind, resident in HBM before the run."""
    build: "G.DeviceRelation | None"
    resident: "G.ResidentBuild | None"
    probe: "G.DeviceRelation"
    cfg: G.RadixConfig
    build_rows: int
    probe_rows: int
    footprint: int  # bytes reserved while in flight
    spec: object = None  # generator spec of (build, probe): its closed-form checksum is the exact result


@dataclass
class Outcome:
    rid: int
    kind: int
    arrival: float
    admitted: float = 0.0
    done: float = 0.0
    out_rows: int = 0
    checksum: tuple | None = None

    @property
    def latency(self):
        return self.done - self.arrival


def _row_bytes(types):
    return 8 + sum(G.TYPE_SIZE[t] for t in types)


def request_pool(ctx, seed=42, n_new=32, n_resident=64, n_resident_probes=16, lo_log2=12, hi_log2=24,
                 probe_lo=4, probe_hi=16, match_lo=0.1, match_hi=1.0, max_cols=3, budget_bytes=None):
    """Request kinds: n_new joins with their own build side and n_resident_probes
    probe sides against n_resident registered build sides. |R| = floor(2^U[lo,hi]),
    |S| = U[probe_lo, probe_hi] |R|, match rate U[match_lo, match_hi], 1..max_cols
    payload columns of {u32, u64, u128} per side; all draws from `seed`."""
    rng = np.random.default_rng(seed)

    def draw_types():
        return tuple(int(t) for t in rng.choice(_PAYLOAD_TYPES, rng.integers(1, max_cols + 1)))

    def footprint(r_rows, s_rows, bt, pt, resident):
        out = s_rows * (8 + sum(G.TYPE_SIZE[t] for t in bt) + sum(G.TYPE_SIZE[t] for t in pt))
        scratch = 2 * s_rows * _row_bytes(pt) + (0 if resident else 2 * r_rows * _row_bytes(bt))
        return out + scratch

    kinds = []
    for i in range(n_new):
        r_rows = int(2.0 ** rng.uniform(lo_log2, hi_log2))
        s_rows = int(rng.uniform(probe_lo, probe_hi) * r_rows)
        bt, pt = draw_types(), draw_types()
        spec = G.gen_spec(seed=int(rng.integers(1, 2**62)), build_rows=r_rows, probe_rows=s_rows,
                          match_rate=float(rng.uniform(match_lo, match_hi)), build_types=bt, probe_types=pt)
        b, p = ctx.generate(spec)
        kinds.append(RequestInput(b, None, p, G.RadixConfig.for_build_rows(r_rows), r_rows, s_rows,
                                  footprint(r_rows, s_rows, bt, pt, False), spec))
    residents = []
    for i in range(n_resident):
        r_rows = int(2.0 ** rng.uniform(lo_log2, hi_log2))
        bt = draw_types()
        rseed = int(rng.integers(1, 2**62))
        spec = G.gen_spec(seed=rseed, build_rows=r_rows, probe_rows=0, build_types=bt, probe_types=())
        b, p = ctx.generate(spec)
        res = ctx.register_build(b, G.RadixConfig.for_build_rows(r_rows))
        b.release()
        p.release()
        residents.append((res, rseed, r_rows, bt))
    for i in range(n_resident_probes):
        res, rseed, r_rows, bt = residents[int(rng.integers(0, n_resident))]
        s_rows = int(rng.uniform(probe_lo, probe_hi) * r_rows)
        pt = draw_types()
        # same seed and build rows as the registered side: the probe keys match it
        spec = G.gen_spec(seed=rseed, build_rows=r_rows, probe_rows=s_rows,
                          match_rate=float(rng.uniform(match_lo, match_hi)), build_types=bt, probe_types=pt)
        b, p = ctx.generate(spec)
        b.release()
        kinds.append(RequestInput(None, res, p, G.RadixConfig.for_build_rows(r_rows), r_rows, s_rows,
                                  footprint(r_rows, s_rows, bt, pt, True), spec))
    return kinds, [r[0] for r in residents], n_new


def expected(ctx, kind: RequestInput):
    """(out_rows, checksum) of the request's exact result, from the generator's closed form."""
    cs = ctx.expected_checksum(kind.spec)
    return cs[0], cs


def run_request(ctx, kind: RequestInput, want_checksum=False):
    if kind.resident is not None:
        r = kind.resident.probe(kind.probe, ctx=ctx)
    else:
        r = ctx.runPHJ(kind.build, kind.probe, kind.cfg)
    try:
        return r.out_rows, (r.checksum() if want_checksum else None)
    finally:
        r.release()


def schema_of(kind: RequestInput):
    b = kind.build.types if kind.build is not None else None
    return (tuple(b) if b is not None else None, tuple(kind.probe.types))


def run_batch(ctx, batch, want_checksum=False):
    """Requests with their own (small) build sides joined in one
    phj_join_batch call; returns (out_rows, checksum) per request."""
    res = ctx.join_batch([(k.build, k.probe) for k in batch], [k.cfg for k in batch])
    try:
        return [(r.out_rows, r.checksum() if want_checksum else None) for r in res]
    finally:
        for r in res:
            r.release()


class JoinServer:
    """W workers (one Context each) behind a FIFO admission queue bounded by HBM bytes.

    Continuous batching: a worker that admits a request with its own build
    side of at most batch_rows build rows also admits every other queued
    request of that kind (small, same schema) that fits the budget, up to
    max_batch, and joins them in one phj_join_batch call."""

    def __init__(self, workers=4, budget_bytes=200 << 30, device=0, batch_rows=1 << 18, max_batch=64):
        self.ctxs = [G.Context(device) for _ in range(workers)]
        self.budget = budget_bytes
        self.batch_rows = batch_rows
        self.max_batch = max_batch
        self.batches = 0           # batched calls made in the last run
        self.batched_requests = 0  # requests they joined

    def close(self):
        for c in self.ctxs:
            c.close()

    def run(self, kinds, schedule, verify_every=0):
        """schedule: list of (arrival_seconds, kind_index). Returns Outcomes.

        Arrivals join a FIFO queue. A worker admits the head of the queue when
        its footprint fits the budget beside the requests in flight (or nothing
        is in flight), runs it, and releases the footprint when it completes."""
        cond = threading.Condition()
        pending = []  # arrived, not yet admitted (FIFO)
        state = {"inflight": 0, "arrived": 0, "done": False}
        outcomes = [Outcome(i, k, t) for i, (t, k) in enumerate(schedule)]
        errors = []
        t0 = time.perf_counter()

        def small(k):
            return k.resident is None and k.build_rows <= self.batch_rows

        def worker(ctx):
            while True:
                with cond:
                    while True:
                        if pending:
                            need = kinds[outcomes[pending[0]].kind].footprint
                            if state["inflight"] == 0 or state["inflight"] + need <= self.budget:
                                i = pending.pop(0)
                                state["inflight"] += need
                                break
                        elif state["done"]:
                            return
                        cond.wait()
                    group = [i]
                    if small(kinds[outcomes[i].kind]):
                        sch = schema_of(kinds[outcomes[i].kind])
                        for j in list(pending):
                            if len(group) >= self.max_batch:
                                break
                            kj = kinds[outcomes[j].kind]
                            if small(kj) and schema_of(kj) == sch and state["inflight"] + kj.footprint <= self.budget:
                                pending.remove(j)
                                state["inflight"] += kj.footprint
                                group.append(j)
                now = time.perf_counter() - t0
                for g in group:
                    outcomes[g].admitted = now
                want = [bool(verify_every) and g % verify_every == 0 for g in group]
                try:
                    if len(group) == 1:
                        outs = [run_request(ctx, kinds[outcomes[i].kind], want[0])]
                    else:
                        outs = run_batch(ctx, [kinds[outcomes[g].kind] for g in group], any(want))
                        with cond:
                            self.batches += 1
                            self.batched_requests += len(group)
                    for g, w, (rows, cs) in zip(group, want, outs):
                        outcomes[g].out_rows, outcomes[g].checksum = rows, (cs if w else None)
                except Exception as e:  # reported by run(); the footprint is still released
                    errors.append((i, repr(e)))
                done = time.perf_counter() - t0
                with cond:
                    for g in group:
                        outcomes[g].done = done
                        state["inflight"] -= kinds[outcomes[g].kind].footprint
                    cond.notify_all()

        self.batches = self.batched_requests = 0
        threads = [threading.Thread(target=worker, args=(c,), daemon=True) for c in self.ctxs]
        for th in threads:
            th.start()
        for i, o in enumerate(outcomes):
            delay = o.arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            with cond:
                pending.append(i)
                cond.notify_all()
        with cond:
            state["done"] = True
            cond.notify_all()
        for th in threads:
            th.join()
        if errors:
            raise RuntimeError(f"{len(errors)} requests failed, first: request {errors[0][0]}: {errors[0][1]}")
        return outcomes


def poisson_schedule(rate, n_requests, n_kinds, seed=7, n_new=None, resident_share=None):
    """Arrival times of a Poisson process of `rate` per second and the kind of
    each request: uniform over the kinds, or, with n_new and resident_share,
    a resident-build kind (index >= n_new) with probability resident_share."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    if n_new is None or resident_share is None or n_new >= n_kinds:
        kinds = rng.integers(0, n_kinds, n_requests)
    else:
        res = rng.random(n_requests) < resident_share
        kinds = np.where(res, rng.integers(n_new, n_kinds, n_requests), rng.integers(0, n_new, n_requests))
    return [(float(a), int(k)) for a, k in zip(t, kinds)]


def summarize(outcomes, kinds):
    lat = np.array([o.latency for o in outcomes])
    span = max(o.done for o in outcomes) - min(o.arrival for o in outcomes)
    tuples = sum(kinds[o.kind].build_rows + kinds[o.kind].probe_rows for o in outcomes)
    return {"requests": len(outcomes), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
            "completed_per_s": len(outcomes) / span, "tuples_per_s": tuples / span, "makespan_s": span}
