"""Benchmark: PHJ-PURE hash join, 2^27 build rows x 2^30 probe rows (uint64 keys),
build payloads u32+u64, probe payloads u64+u128, 50% of probe rows match,
radix 2 passes x 256 partitions, load factor 0.5, on one MI355X.

One step = one full runPHJ (build shuffle, probe shuffle, per-leaf build and
probe, output materialised in HBM) over inputs already resident in HBM.
The line's "batched" object measures 4096 small joins (|R_i| = 2^U[10,20],
|S_i| = 8|R_i|) as one phj_join_batch call against one runPHJ call per join;
"skew" times PHJ-PURE on a Zipf-skewed join (2^22 keys with repeated build
rows vs 2^27 probes) and checks CHJ and PHJ agree with it; "serving" runs
open-loop Poisson requests through the serving loop (phj-bench_amd/serve.py)
at several offered rates (p50/p99 latency).
Prints one JSON line.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--scale-log2 0] [--no-cpu-baseline] [--batch-joins 4096]
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "phj-bench_amd"))

import phj_gpu as G  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "join input tuples/s & % HBM peak, 2^27⋈2^30 u64; batched vs plain speedup"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def widths(types):
    return sum(G.TYPE_SIZE[t] for t in types)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_threads():
    """Threads the CPU baseline uses: the process's CPU share (OMP_NUM_THREADS
    when set, as on a shared GPU host), else the cores it may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(scale_log2, threads):
    """The compiled reference (oracle/_ref) on this host's cores, timed by its
    own e2e_wall_ms (README.md:144-216), best of 2 after a warm-up:
    - value: its runPHJ on a bounded sample of the headline workload
      (2^22 x 2^25 rows, same distribution, schema and pass_bits {8,8});
    - config1: CHJ and PHJ-PURE on 2^20 x 2^24 in full (distinct uniform u64
      build keys in [1, 2^63), every probe row matching once, u64 payloads,
      pass_bits {8}), at `threads` threads and at 1 thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    if not O.ref_available():
        return None

    def best_ms(scheme, b, p, bits, t):
        O.ref_run(scheme, b, p, bits, threads=t, want_checksum=False)  # warm-up
        return min(O.ref_run(scheme, b, p, bits, threads=t, want_checksum=False).e2e_wall_ms for _ in range(2))

    rb, rp = 22 - scale_log2, 25 - scale_log2
    spec = O.make_spec(seed=42, build_rows=1 << rb, probe_rows=1 << rp, distribution=0, match_rate=0.5,
                       build_types=(G.U32, G.U64), probe_types=(G.U64, G.U128))
    b, p = O.gen(spec, threads=threads)
    ms = best_ms(3, b, p, (8, 8), threads)
    log(f"CPU baseline: reference runPHJ 2^{rb} x 2^{rp} on {threads} threads: {ms:.1f} ms")
    del b, p
    c1 = O.make_spec(seed=42, build_rows=1 << 20, probe_rows=1 << 24, distribution=0, match_rate=1.0,
                     build_types=(G.U64,), probe_types=(G.U64,))
    b1, p1 = O.gen(c1, threads=threads)
    config1 = {"workload": "2^20 x 2^24, distinct uniform u64 build keys, 100% match, u64 payloads, pass_bits {8}",
               "tuples": b1.rows + p1.rows}
    for name, scheme in (("chj", 1), ("phj_pure", 3)):
        for t in sorted({threads, 1}, reverse=True):
            m = best_ms(scheme, b1, p1, (8,), t)
            config1[f"{name}_{t}t_ms"] = round(m, 2)
            config1[f"{name}_{t}t_tuples_per_s"] = (b1.rows + p1.rows) / (m / 1e3)
            log(f"CPU baseline: reference {name} config 1 on {t} threads: {m:.1f} ms")
    return {"value": ((1 << rb) + (1 << rp)) / (ms / 1e3),
            "unit": "tuples/s", "cores": threads, "kind": "reference", "cpu_model": cpu_model(),
            "sample": f"reference runPHJ (oracle/_ref, its own e2e_wall_ms), 2^{rb} x 2^{rp} rows of the same "
                      f"distribution and schema, pass_bits {{8,8}}, best of 2 after 1 warm-up",
            "config1": config1}


def cpu_baseline_child(scale_log2, threads):
    """cpu_baseline in a child process (it runs the compiled reference, whose
    faults must not end the benchmark): its JSON, or None with the failure
    logged. One retry."""
    for attempt in (1, 2):
        log(f"CPU baseline: compiled reference on {threads} threads (child process, attempt {attempt})")
        try:
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--scale-log2",
                                str(scale_log2), "--cpu-threads", str(threads)], capture_output=True, text=True,
                               timeout=600)
        except subprocess.TimeoutExpired:
            log("CPU baseline: child timed out")
            continue
        for line in r.stderr.splitlines():
            if line.startswith("[bench]"):
                print(line, file=sys.stderr, flush=True)
        if r.returncode == 0 and r.stdout.strip():
            return json.loads(r.stdout.strip().splitlines()[-1])
        log(f"CPU baseline: child exited with {r.returncode}")
    return None


def batched_bench(ctx, n_joins, seed=42, plain=True):
    """4096 independent uniform joins (|R_i| = floor(2^U[10,20]) distinct keys,
    |S_i| = 8 |R_i|, 75% match, u64 payload each side), back to back in HBM:
    one phj_join_batch call (median of 3) against one runPHJ call per join.
    Per-join checksums of the two must agree."""
    specs = G.batched_join_specs(seed=seed, n_joins=n_joins)
    build, probe, boff, poff = ctx.generate_batch(specs)
    pairs = [(build.slice(int(boff[i]), int(boff[i + 1])), probe.slice(int(poff[i]), int(poff[i + 1])))
             for i in range(n_joins)]
    cfgs = [G.RadixConfig.for_build_rows(b.rows) for b, _ in pairs]
    R, S = build.rows, probe.rows
    log(f"batch: {n_joins} joins, {R} build + {S} probe rows")

    def run_batched():
        return ctx.join_batch(pairs, cfgs)

    def run_plain():
        return [ctx.runPHJ(b, p, cfg) for (b, p), cfg in zip(pairs, cfgs)]

    def timed(run):
        ctx.synchronize()
        t0 = time.perf_counter()
        res = run()
        ctx.synchronize()
        return time.perf_counter() - t0, res

    # One untimed run of the whole workload per path first: the context's
    # device pool then holds the memory both paths need (fresh device memory
    # is cleared by the driver when first mapped, which is not join work).
    for run in ((run_batched, run_plain) if plain else (run_batched,)):
        for r in run():
            r.release()

    def median_batched():
        ts = []
        for _ in range(3):
            t, res = timed(run_batched)
            ts.append(t)
            got = [r.checksum() for r in res]
            rows = sum(r.out_rows for r in res)
            for r in res:
                r.release()
        return sorted(ts)[1], got, rows

    t_default, _, _ = median_batched()
    # the autotuner (untimed) picks a leaf size per size class; the batch is
    # then timed with the tuned table
    t0 = time.perf_counter()
    table = ctx.batch_autotune(pairs, cfgs)
    t_tune = time.perf_counter() - t0
    log(f"batch: autotune {t_tune:.1f} s: " + ", ".join(f"2^{c}:{v[0]}" for c, v in sorted(table.items())))
    t_batch, got, out_rows = median_batched()
    if plain:
        t_plain, res_plain = timed(run_plain)
        want = [r.checksum() for r in res_plain]
        for r in res_plain:
            r.release()
    else:
        t_plain, want = float("nan"), [ctx.expected_checksum(sp) for sp in specs]
    build.release()
    probe.release()
    parity = got == want
    log(f"batch: batched {t_batch * 1e3:.1f} ms (untuned {t_default * 1e3:.1f} ms), plain {t_plain * 1e3:.1f} ms, "
        f"per-join parity {parity}")
    # algorithmic bytes: every input row read once, every output row written once
    b_alg = R * 16 + S * 16 + out_rows * 24
    return {"joins": n_joins, "build_rows": R, "probe_rows": S, "out_rows": out_rows,
            "batched_ms": t_batch * 1e3, "batched_untuned_ms": t_default * 1e3, "plain_ms": t_plain * 1e3,
            "speedup": t_plain / t_batch, "roofline_frac": b_alg / t_batch / HBM_PEAK,
            "autotune": {f"2^{c}": {"leaf_rows": v[0], "best_ms": round(v[1], 3)} for c, v in sorted(table.items())},
            "autotune_s": t_tune,
            "batched_tuples_per_s": (R + S) / t_batch, "plain_tuples_per_s": (R + S) / t_plain,
            "parity_per_join": parity,
            "workload": f"{n_joins} joins, |R_i| = floor(2^U[10,20]), |S_i| = 8|R_i|, 75% match, u64 payloads"}


def skew_bench(reps=2, scale_log2=0):
    """2^22 distinct build keys, each repeated k ~ Zipf(1.2) on [1,64] times and
    shuffled (u8+u16+u64 payload); 2^27 probe rows, 90% drawing a key by
    Zipf(0.8) rank, 10% missing (u64 payload). PHJ-PURE timed; CHJ and PHJ
    (best-effort, M = 32 MiB) must give the closed-form checksum."""
    ctx = G.Context(0)
    spec = G.gen_spec(seed=42, build_rows=1 << (22 - scale_log2), probe_rows=1 << (27 - scale_log2),
                      distribution=G.DIST_ZIPF, match_rate=0.9,
                      build_types=(G.U8, G.U16, G.U64), probe_types=(G.U64,), zipf_max_dup=64, zipf_build_s=1.2,
                      zipf_probe_s=0.8)
    build, probe = ctx.generate(spec)
    log(f"skew: generated {build.rows} x {probe.rows} rows")
    cfg = G.RadixConfig([8, 8])
    ctx.runPHJ(build, probe, cfg).release()
    log("skew: warm-up PHJ-PURE done")
    best, res = None, None
    for _ in range(reps):
        if res is not None:
            res.release()
        ctx.synchronize()
        t0 = time.perf_counter()
        res = ctx.runPHJ(build, probe, cfg)
        ms = (time.perf_counter() - t0) * 1e3
        best = ms if best is None else min(best, ms)
    want = ctx.expected_checksum(spec)  # closed form: every hit pairs with every build row of its key
    sums = {"phj_pure": res.checksum() == want}
    out_rows = res.out_rows
    res.release()
    log(f"skew: PHJ-PURE best {best:.2f} ms")
    r = ctx.runCHJ(build, probe)
    sums["chj"] = r.checksum() == want
    log(f"skew: CHJ e2e {r.e2e_wall_ms:.2f} ms")
    r.release()
    # M = 32 MiB: the device runs one probe scheduler, so the per-worker budget
    # is the whole budget (src/PHJBep.cpp:186-190 with one worker)
    r = ctx.runPhjBep(build, probe, cfg, bep_budget_mib=32)
    sums["phj_bep"] = r.checksum() == want
    bep = {"budget_mib": r.bep_budget_mib, "evictions": r.bep_evictions, "refinements": r.bep_refinements,
           "peak_bytes": r.bep_peak_bytes, "e2e_ms": r.e2e_wall_ms,
           "probe_shuffle_ms": r.probe_shuffle.wall_ms, "probe_ms": r.probe.wall_ms,
           "eviction_overhead_ms": r.eviction_overhead.wall_ms, "build_ms": r.build.wall_ms}
    log(f"skew: PHJ (32 MiB) e2e {r.e2e_wall_ms:.1f} ms, {r.bep_refinements} refinements, "
        f"{r.bep_evictions} evictions, peak {r.bep_peak_bytes >> 20} MiB")
    r.release()
    R, S = build.rows, probe.rows
    build.release()
    probe.release()
    ctx.close()
    log(f"skew: {R} x {S} rows, {out_rows} out, PHJ-PURE {best:.2f} ms, schemes agree {sums}")
    return {"build_rows": R, "probe_rows": S, "out_rows": out_rows, "phj_pure_ms": best,
            "tuples_per_s": (R + S) / (best / 1e3), "schemes_agree": all(sums.values()), "agree": sums,
            "bep": bep, "workload": "2^22 keys x Zipf(1.2) dups on [1,64] (u8+u16+u64) vs 2^27 probes, 90% "
                                    "Zipf(0.8)-ranked hits (u64), pass_bits {8,8}"}


def serving_bench(n_requests, rates, workers=4, budget_gb=200, scale_log2=0):
    """Open-loop Poisson requests (|R| = 2^U[12,24], |S| = U[4,16] |R|, match
    U[0.1,1], 1-3 payload columns of u32/u64/u128 per side; 30% probe one of 64
    resident build sides) through the serving loop at each offered rate. Every
    20th request's checksum, and every request kind run alone, is compared with
    the closed form of its generator spec. Small requests are batched
    continuously (serve.JoinServer)."""
    import serve

    ctx = G.Context(0)
    kinds, residents, n_new = serve.request_pool(ctx, seed=42, n_new=32, n_resident=64, n_resident_probes=16,
                                                 hi_log2=24 - scale_log2)
    log(f"serve: {len(kinds)} request kinds, {len(residents)} resident build sides ready")
    alone = [serve.run_request(ctx, k, want_checksum=True) for k in kinds]
    log("serve: every request kind run alone")
    srv = serve.JoinServer(workers=workers, budget_bytes=budget_gb << 30)
    # warm-up: every worker's memory pool grows to its working size
    srv.run(kinds, serve.poisson_schedule(200, 4 * len(kinds), len(kinds), seed=1))
    log("serve: warm-up done")
    exact = [serve.expected(ctx, k) for k in kinds]
    parity = alone == exact
    points = []
    for rate in rates:
        sched = serve.poisson_schedule(rate, n_requests, len(kinds), seed=rate, n_new=n_new, resident_share=0.3)
        outs = srv.run(kinds, sched, verify_every=20)
        parity = parity and all(o.checksum is None or (o.out_rows, o.checksum) == exact[o.kind] for o in outs)
        sm = serve.summarize(outs, kinds)
        sm["offered_per_s"] = rate
        sm["batched_calls"] = srv.batches
        sm["batched_requests"] = srv.batched_requests
        points.append({k: (round(v, 3) if isinstance(v, float) else v) for k, v in sm.items()})
        log(f"serve {rate}/s: p50 {sm['p50_ms']:.2f} ms p99 {sm['p99_ms']:.2f} ms, "
            f"{sm['completed_per_s']:.0f} req/s, {sm['tuples_per_s']:.3g} tuples/s")
    srv.close()
    for k in kinds:
        for rel in (k.build, k.probe):
            if rel is not None:
                rel.release()
    for r in residents:
        r.release()
    ctx.close()
    sat = max(points, key=lambda q: q["completed_per_s"])
    return {"workers": workers, "budget_gb": budget_gb, "requests_per_point": n_requests, "points": points,
            "saturation_req_per_s": sat["completed_per_s"], "saturation_tuples_per_s": sat["tuples_per_s"],
            "parity_alone_vs_mix": parity, "parity": "every 20th request, and every kind alone, against the "
                                                   "generator's closed-form checksum",
            "workload": "Poisson arrivals; |R| = 2^U[12,24], |S| = U[4,16]|R|, match U[0.1,1], 1-3 payload columns "
                        "of u32/u64/u128 per side, 30% of requests probe one of 64 resident build sides"}


def headline(args, tuning):
    """The headline join (config 2): warm-up, timed steps, checksum, roofline."""
    rb, rp = 27 - args.scale_log2, 30 - args.scale_log2
    R, S = 1 << rb, 1 << rp
    btypes, ptypes = (G.U32, G.U64), (G.U64, G.U128)
    pass_bits = [8, 8]
    spec_kw = dict(seed=42, build_rows=R, probe_rows=S, distribution=G.DIST_UNIFORM, match_rate=0.5,
                   build_types=btypes, probe_types=ptypes)
    spec = G.gen_spec(**spec_kw)
    ctx = G.Context(0)
    log(f"generating build 2^{rb} and probe 2^{rp} rows on the device")
    build, probe = ctx.generate(spec)
    cfg = G.RadixConfig(pass_bits)

    for w in range(args.warmup):
        r = ctx.runPHJ(build, probe, cfg, tuning=tuning)
        log(f"warmup {w}: e2e {r.e2e_wall_ms:.2f} ms out_rows {r.out_rows}")
        r.release()

    ctx.set_kernel_timing(True)
    ctx.reset_kernel_times()
    ctx.synchronize()
    e2e, phases = [], []
    last = None
    t0 = time.perf_counter()
    for k in range(args.steps):
        if last is not None:
            last.release()
        last = ctx.runPHJ(build, probe, cfg, tuning=tuning)
        e2e.append(last.e2e_wall_ms)
        phases.append((last.build_shuffle.wall_ms, last.probe_shuffle.wall_ms, last.build.wall_ms, last.probe.wall_ms))
    ctx.synchronize()
    t1 = time.perf_counter()
    ktimes = ctx.kernel_times()
    ctx.set_kernel_timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step (wall), e2e events {['%.2f' % x for x in e2e]}")
    marked = ctx.placement_stats()
    log(f"scatter placement watch: {marked[0]} output blocks marked slow, {marked[1] >> 20} MiB marked now")

    if args.calib_copy_mib:
        ms = ctx.copy_bench(args.calib_copy_mib << 20, reps=1)
        log(f"calibration copy of {args.calib_copy_mib} MiB: {ms:.3f} ms")

    out_rows = last.out_rows
    got = last.checksum()
    want = ctx.expected_checksum(spec)
    parity = got == want
    log(f"checksum gpu={got} closed-form={want} parity={parity}")
    last.release()

    wR, wS = 8 + widths(btypes), 8 + widths(ptypes)
    wO = 8 + widths(btypes) + widths(ptypes)
    b_alg = R * wR + S * wS + out_rows * wO
    tuples_per_s = (R + S) / (ms_per_step / 1e3)
    best_e2e = min(e2e)

    # Dominant kernel and its algorithmic bytes per launch (mean over the
    # launches of one step). Per step: 2 passes x 2 relations of k_hist and
    # k_scatter, then the fused leaf join (or a probe pass and an emit pass).
    n = args.steps
    passes = 2
    per_step_bytes = {
        "radix_scatter": passes * 2 * (R * wR + S * wS),  # every pass reads and writes each row once
        # the first pass reads the keys; the second the digit byte column the first scatter wrote
        "radix_hist": (8 + 1) * (R + S),
        # fused leaf join: both sides' rows read once, output written once
        "leaf_join": R * wR + S * wS + out_rows * wO,
        # two-phase leaf join (repeated build keys or PHJ_TUNE_TWO_PHASE_JOIN)
        "leaf_probe": 8 * (R + S),                         # keys of both sides read once
        # payloads of both sides read once, output written once
        "leaf_emit": R * (wR - 8) + S * (wS - 8) + out_rows * wO,
    }
    dom = max(ktimes, key=lambda k: ktimes[k][0]) if ktimes else None
    roofline = None
    if dom in per_step_bytes:
        tot_ms, launches = ktimes[dom]
        avg_ms = tot_ms / launches
        bytes_per_launch = per_step_bytes[dom] * n / launches
        achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", "traffic.json")
        if os.path.exists(tf) and args.scale_log2 == 0:
            traffic = json.load(open(tf)).get(dom)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                    "frac": round(achieved / (HBM_PEAK / 1e9), 4), "traffic": traffic, "kernel": dom,
                    "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": int(bytes_per_launch)}

    build.release()
    probe.release()
    if roofline is not None and args.scale_log2 == 0:
        # the part's measured streaming rates (4 GiB buffers), beside the 8 TB/s peak
        rates = ctx.stream_bench(4 << 30, reps=5)
        roofline["measured_stream_GBps"] = {k: round(v / 1e9, 1) for k, v in rates.items()}
        roofline["frac_of_copy"] = round(roofline["achieved"] / (rates["copy"] / 1e9), 4)
        log("streaming rates: " + ", ".join(f"{k} {v / 1e9:.0f} GB/s" for k, v in rates.items()))
    ctx.close()  # each workload below runs in a context of its own, on a device with nothing cached
    return dict(R=R, S=S, rb=rb, rp=rp, pass_bits=pass_bits, ms_per_step=ms_per_step, tuples_per_s=tuples_per_s,
                best_e2e=best_e2e, b_alg=b_alg, out_rows=out_rows, parity=parity, phases=phases, ktimes=ktimes,
                marked=marked,
                roofline=roofline, n=n)


def layout_probes(scale_log2=0):
    """The measurements behind the partitioning layout (phj_layout_probe):
    random 16-byte lookups by table size (where a table must live for a pass
    to be saved), a buffer written and read back at once by scratch size
    (whether the caches absorb a fused pass's intermediate), and one radix
    pass's memory pattern over 2^30 rows of 32 bytes into column or row
    outputs and into 256..4,096 partitions (the cost of a wider single pass)."""
    ctx = G.Context(0)
    reads = 1 << (28 - scale_log2)
    rr = {}
    for mib in (1, 2, 4, 8, 16, 32, 64, 128, 256, 1024, 8192):
        ms = ctx.layout_probe(G.Context.PROBE_RANDOM_READ, mib << 20, reads, reps=3)
        rr[f"{mib} MiB"] = round(reads / (ms * 1e-3) / 1e9, 1)
    log("random 16-byte reads, G/s by table size: " + ", ".join(f"{k} {v}" for k, v in rr.items()))
    nb = 16 << (30 - scale_log2)
    sc = {}
    for mib in (16, 32, 64, 128, 256, 512, nb >> 20):
        sc[f"{mib} MiB"] = round(ctx.layout_probe(G.Context.PROBE_SCRATCH_COPY, nb, mib << 20, reps=3), 3)
    base = sc[f"{nb >> 20} MiB"]
    log("copy of %d GiB through a scratch, ms by scratch size: " % (nb >> 30) +
        ", ".join(f"{k} {v}" for k, v in sc.items()) + f" (the last holds the whole buffer: no reuse)")
    pb = 32 << (30 - scale_log2)
    passes = {}
    for parts in (256, 512, 1024, 2048, 4096):
        passes[f"columns {parts}"] = round(ctx.layout_probe(G.Context.PROBE_PASS_COLUMNS, pb, parts, reps=3), 3)
    passes["rows 256"] = round(ctx.layout_probe(G.Context.PROBE_PASS_ROWS, pb, 256, reps=3), 3)
    log("one pass over %d rows of 32 bytes, ms: " % (pb // 32) + ", ".join(f"{k} {v}" for k, v in passes.items()))
    ctx.close()
    # one pass into fresh output columns: four sets of separately allocated
    # columns, then four sets each allocated as one block (columns back to back)
    pctx = G.Context(0)
    placement = {"separate": [round(v, 3) for v in pctx.placement_probe(pb // 32, 4, reps=3)],
                 "one_block": [round(v, 3) for v in pctx.placement_probe(pb // 32, 4, reps=3, gap=0)]}
    pctx.close()
    log("the same pass (columns, 256) into 4 freshly allocated sets, ms: separate columns " +
        ", ".join(map(str, placement["separate"])) + "; one block per set " + ", ".join(map(str, placement["one_block"])))
    return {"random_read_G_per_s": rr, "reads_per_size": reads,
            "scratch_copy_ms": sc, "scratch_copy_bytes": nb, "scratch_copy_no_reuse_ms": base,
            "pass_ms": passes, "pass_rows": pb // 32, "placement_sets_ms": placement,
            "workload": "phj_layout_probe: random 16-byte reads by table size; a copy through a reused scratch; "
                        "one radix pass's memory pattern (32-byte rows, 4,096-row tiles) into columns or rows"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale-log2", type=int, default=0, help="shrink both sides by 2^k (testing only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one plain device copy of this size (HBM counter calibration under a profiler)")
    ap.add_argument("--tune", default="", help="phj_tuning overrides, e.g. flags=12,join_threads=1024")
    ap.add_argument("--batch-joins", type=int, default=4096, help="joins of the batched-join measurement (0: skip)")
    ap.add_argument("--batch-no-plain", action="store_true",
                    help="skip the per-join runPHJ comparison of the batch (profiling the batch alone)")
    ap.add_argument("--serve-requests", type=int, default=400, help="requests per offered rate of the serving "
                                                                     "measurement (0: skip)")
    ap.add_argument("--serve-rates", default="250,500,1000,2000,5000,10000,20000")
    ap.add_argument("--no-skew", action="store_true", help="skip the skewed-join measurement")
    ap.add_argument("--workloads", default="main,batch,skew,serve",
                    help="comma-separated subset of main (the headline join), batch, skew, serve, probes (the "
                         "layout probes, not in the default set); a run without main prints value null (for "
                         "profiling one workload)")
    ap.add_argument("--cpu-baseline-only", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline_only:  # the child process of cpu_baseline_child; touches no GPU
        print(json.dumps(cpu_baseline(args.scale_log2, args.cpu_threads or host_threads())), flush=True)
        return
    wl = set(w for w in args.workloads.split(",") if w)
    if "batch" not in wl:
        args.batch_joins = 0
    if "serve" not in wl:
        args.serve_requests = 0
    if "skew" not in wl:
        args.no_skew = True
    tuning = {k: int(v) for k, v in (kv.split("=") for kv in args.tune.split(",") if kv)}
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the join on a single MI355X; "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)

    h = headline(args, tuning) if "main" in wl and args.steps > 0 else None
    batched = None
    if args.batch_joins > 0:
        bctx = G.Context(0)
        batched = batched_bench(bctx, args.batch_joins, plain=not args.batch_no_plain)
        bctx.close()
    skew = skew_bench(scale_log2=args.scale_log2) if not args.no_skew else None
    serving = None
    if args.serve_requests > 0:
        serving = serving_bench(args.serve_requests, [int(r) for r in args.serve_rates.split(",") if r],
                                scale_log2=args.scale_log2)

    probes = layout_probes(args.scale_log2) if "probes" in wl else None

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_threads()
        cpu = cpu_baseline_child(args.scale_log2, threads)

    hv = h or {}
    line = {
        "metric": METRIC,
        "value": hv.get("tuples_per_s"),
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hv.get("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (device-generated, seed 42)",
        "config": None if h is None else {
            "workload": f"PHJ-PURE 2^{h['rb']} x 2^{h['rp']}, build u32+u64, probe u64+u128, 50% match, "
                        f"pass_bits {{8,8}}, load factor 0.5; leaf tables in LDS as u64 keys + u16 rows "
                        f"(~10 B per slot; the 16-byte (key, RowRefCell) cells are the HBM tables' layout)",
            "build_rows": h["R"], "probe_rows": h["S"], "pass_bits": h["pass_bits"], "match_rate": 0.5},
        "roofline": hv.get("roofline"),
        "cpu_baseline": cpu,
        "e2e_ms_best": hv.get("best_e2e"),
        "e2e_roofline_frac": None if h is None else h["b_alg"] / (h["best_e2e"] / 1e3) / HBM_PEAK,
        "step_roofline_frac": None if h is None else h["b_alg"] / (h["ms_per_step"] / 1e3) / HBM_PEAK,
        "algorithmic_bytes": hv.get("b_alg"),
        "out_rows": hv.get("out_rows"),
        "parity_checksum": hv.get("parity"),
        "phase_ms_last": None if h is None else dict(zip(("build_shuffle", "probe_shuffle", "build", "probe"),
                                                          h["phases"][-1])),
        "kernel_ms_per_step": None if h is None else {k: round(v[0] / h["n"], 4) for k, v in sorted(h["ktimes"].items())},
        "placement_marked": None if h is None else {"blocks": h["marked"][0], "bytes": h["marked"][1]},
        "batched": batched,
        "skew": skew,
        "serving": serving,
    }
    if probes is not None:
        line["layout_probes"] = probes
    print(json.dumps(line), flush=True)
    if ((h is not None and not h["parity"]) or (batched is not None and not batched["parity_per_join"])
            or (serving is not None and not serving["parity_alone_vs_mix"])
            or (skew is not None and not skew["schemes_agree"])):
        sys.exit(1)


if __name__ == "__main__":
    main()
