"""Benchmark: exact self-join of 1,000,000 sets at Jaccard >= 0.8 on one MI355X.

Workload (the headline configuration): 1M sets, sizes round(lognormal(median
10, sigma 0.6)) clipped to [2, 200], tokens Zipf(1.1) over 1M ranks re-ranked
by ascending frequency, sets sorted by size, seed 1; Jaccard threshold 0.8
with the reference's double-precision bounds; prefix-filter (fgss)
technique.  A step is one full join from the device-resident CSR collection:
index build + candidate generation + verification, result pairs left in HBM
(their count read back); `value_with_d2h` in the extra fields is the same
step with every result triple copied into host memory.  Prints one JSON line on stdout; progress goes to
stderr.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, published


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_bytes(st: dict, n_sets: int, universe: int, n_tokens: int) -> dict:
    """Algorithmic HBM bytes of the join's stages, per unit of work counted
    by the join itself:

    * index build, per index entry: B_idx = 8 P_idx + 8 n + 8 (U + 1) (read
      the prefix token and write the posting's set id; read each set's start
      and size; write each token's count and offset);
    * signatures, per token: B_sig = 4 N + 8 n + (b / 8) n;
    * candidate generation, per posting visited: B_cand = 12 P_probe + 4 E +
      12 C (each probe-prefix token with the two offsets around it, 4 B per
      posting inside the length window, a 12 B candidate record);
    * verification, per candidate: B_ver = 12 C + 4 sum_cand |s| +
      4 sum_{r with a candidate} |r| + 12 K (the candidate record, both
      sets' tokens, one result triple per pair)."""
    P, E, C, K = st["probe_prefix_tokens"], st["posting_visits"], st["candidates"], st["pairs"]
    b = st["signature_bits"]
    return {
        "index": 8 * st["index_entries"] + 8 * n_sets + 8 * (universe + 1),
        "signatures": 4 * n_tokens + 8 * n_sets + (b // 8) * n_sets if b else 0,
        "candidates": 12 * P + 4 * E + 12 * C,
        "verify": 12 * C + 4 * st["candidate_tokens"] + 4 * st["candidate_probe_tokens"] + 12 * K,
    }


def token_join_design_bytes(st: dict) -> int:
    """What the token-major join kernels read and write by design, each
    datum counted once: every entry of the token-major index (the probe
    prefixes: key, set id, size/position word, 12 B) staged by the filter
    with its set's signature (W words of 8 B); per surviving candidate a
    16 B queue record written and read back, the two sets' starts and sizes
    (16 B) and tokens; 12 B per result triple.  Staging windows that reach
    back into a long posting list and the second pass over heavy tiles
    re-read entries beyond this."""
    P, C, K = st["probe_prefix_tokens"], st["candidates"], st["pairs"]
    w = st["signature_bits"] // 64
    return ((12 + 8 * w) * P + 48 * C + 4 * st["candidate_tokens"] + 4 * st["candidate_probe_tokens"]
            + 12 * K)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_coprocess(ctx, dI, dP, name, threads, want, reps=3) -> dict:
    """The join shared by the device and `threads` host threads (0: the
    library's default, the usable CPUs less one) through the work-sharing
    queue, forced to share (the gate off), against the device alone and
    against the call with the library's break-even gate on (its default:
    the device alone unless sharing measured faster); all timed as wall
    clock around the call, median of `reps`."""
    alone, shared, gated, infos, ginfos = [], [], [], [], []
    for _ in range(reps + 1):
        a = time.perf_counter()
        assert ctx.join(dI, dP, algo="fgss", count_only=True) == want
        alone.append((time.perf_counter() - a) * 1e3)
        a = time.perf_counter()
        c, info = ctx.join_coprocess(dI, dP, cpu_threads=threads, count_only=True, min_device_ms=-1)
        shared.append((time.perf_counter() - a) * 1e3)
        assert c == want
        infos.append(info)
        a = time.perf_counter()
        c, ginfo = ctx.join_coprocess(dI, dP, cpu_threads=threads, count_only=True)
        gated.append((time.perf_counter() - a) * 1e3)
        assert c == want
        ginfos.append(ginfo)
    alone, shared, infos, gated, ginfos = alone[1:], shared[1:], infos[1:], gated[1:], ginfos[1:]
    k = sorted(range(reps), key=lambda i: shared[i])[reps // 2]
    info = infos[k]
    n = info["device_probes"] + info["host_probes"]
    res = {f"{name}_device_alone_ms": round(statistics.median(alone), 3),
           f"{name}_coprocess_ms": round(shared[k], 3),
           f"{name}_coprocess_host_threads": int(info["host_threads"]),
           f"{name}_coprocess_host_probe_share": round(info["host_probes"] / max(n, 1), 5),
           f"{name}_coprocess_host_pairs": int(info["host_pairs"]),
           f"{name}_coprocess_device_claims": int(info["device_claims"]),
           f"{name}_coprocess_host_claims": int(info["host_claims"]),
           f"{name}_coprocess_index_copy_ms": round(info["copy_ms"], 3),
           f"{name}_coprocess_device_busy_ms": round(info["device_ms"], 3),
           f"{name}_coprocess_host_busy_ms": round(info["host_ms"], 3),
           f"{name}_coprocess_gated_ms": round(statistics.median(gated), 3),
           f"{name}_coprocess_gated_shared_calls": sum(1 - g["gated"] for g in ginfos)}
    log(f"{name}: device alone {res[f'{name}_device_alone_ms']} ms, shared with {info['host_threads']} host "
        f"threads {res[f'{name}_coprocess_ms']} ms (host took {info['host_probes']} of {n} probes, "
        f"{info['host_pairs']} pairs; index copy {info['copy_ms']:.3f} ms); with the break-even gate "
        f"{res[f'{name}_coprocess_gated_ms']} ms ({res[f'{name}_coprocess_gated_shared_calls']} of {reps} calls "
        f"shared)")
    return res


def bench_policies(ctx, d, want, threads) -> dict:
    """The dichotomy's policy pairs (src/dichotomy.cu) at the reference's
    default --split 50 and the queue's cooperative (hybrid) device consumer
    on the headline join: wall time of one call each, after a warm-up."""
    res = {}
    runs = (("cpu_gpu", dict(mode="static", policy="cpu,gpu", algo="fgss", cpu_share=0.5)),
            ("gpu_cpu", dict(mode="static", policy="gpu,cpu", algo="fgss", cpu_share=0.5)),
            ("cpu_hybrid", dict(mode="static", policy="cpu,gpu", algo="hybrid", cpu_share=0.5)),
            ("hybrid_cpu", dict(mode="static", policy="gpu,cpu", algo="hybrid", cpu_share=0.5)),
            ("queue_cooperative", dict(mode="queue", algo="hybrid", min_device_ms=-1)))
    for name, kw in runs:
        tune = None
        if kw["algo"] == "hybrid":
            import hyset_amd as H
            tune = H.Tune(host_threads=max(1, threads // 2))
        ms = []
        for _ in range(2):
            a = time.perf_counter()
            c, info = ctx.join_coprocess(d, cpu_threads=max(1, threads - (threads // 2 if tune else 0)), tune=tune,
                                         count_only=True, **kw)
            ms.append((time.perf_counter() - a) * 1e3)
            assert c == want
        res[f"policy_{name}_ms"] = round(ms[-1], 2)
        res[f"policy_{name}_host_probes"] = int(info["host_probes"])
        log(f"policy {name}: {ms[-1]:.2f} ms (host {info['host_probes']} probes, {info['host_pairs']} pairs; "
            f"device {info['device_probes']} probes, {info['device_pairs']} pairs)")
    return res


def bench_partition_pairs(ctx, H, d, coll, t, want, threads) -> dict:
    """cost::run over partition pairs of the headline collection: the
    analytic model fitted to measured pair timings (every technique, bitmap
    at 64/128/256 bits) of 2, 3 and 6 even partitions (blocks of 10,000
    sets, the reference's default), then the plan of 4 partitions run by
    hs_run_partition_pairs (CPU pairs on host threads), against each
    technique alone over the same pairs and the whole join in one call."""
    from hyset_amd import split as S
    trains = [H.partition_even(coll.n, 10_000, p) for p in (2, 3, 6)]
    a = time.perf_counter()
    F, T, bt, counts, _ = S.measure_partition_pairs(ctx, d, trains, t, cpu_threads=1)
    meas_s = time.perf_counter() - a
    model = S.AnalyticModel().fit(F, T, bitmap_times=bt)
    parts = H.partition_even(coll.n, 10_000, 4)
    plan = S.partition_plan(ctx, d, parts, model, t)

    def wall(p):
        ms = []
        for _ in range(3):
            n, _, _, w = ctx.run_partition_pairs(d, p, cpu_threads=threads, count_only=True)
            assert n == want
            ms.append(w)
        return round(statistics.median(ms), 3)

    res = {"partition_pairs": {
        "train_pairs": int(F.shape[0]), "measure_s": round(meas_s, 1),
        "plan": {S.TECH_NAMES[k]: sum(1 for *_, x in plan if x == k) for k in range(4)},
        "plan_ms": wall(plan),
        "fixed_ms": {S.TECH_NAMES[k]: wall([(l, r, k) for l, r, _ in plan]) for k in range(4)},
        "whole_join_ms": None,
        "pairs_4_partitions": len(plan)}}
    ms = []
    for _ in range(3):
        a = time.perf_counter()
        assert ctx.join(d, algo="fgss", count_only=True) == want
        ms.append((time.perf_counter() - a) * 1e3)
    res["partition_pairs"]["whole_join_ms"] = round(statistics.median(ms), 3)
    log(f"partition pairs: {res['partition_pairs']}")
    return res


def bench_long_rs(ctx, H) -> dict:
    """The long-set R-S workload: R = S = 200,000 sets of 100-700 tokens over
    a 50,000-token Zipf(0.8) universe, 20% of S edited copies of R sets,
    cosine 0.8: fgss (index over S, probes R) and the bitmap technique with
    1024-bit signatures."""
    from hyset_amd import workloads as W
    a = time.perf_counter()
    R, S = W.long_rs()
    log(f"long R-S: {R.n} x {S.n} sets, {R.tokens.shape[0] + S.tokens.shape[0]} tokens, "
        f"generated in {time.perf_counter() - a:.1f}s")
    dR, dS = ctx.upload(R), ctx.upload(S)
    ctx.set_threshold(0.8, "cosine")
    c = ctx.join(dS, dR, algo="fgss", count_only=True)
    ms = []
    for _ in range(3):
        assert ctx.join(dS, dR, algo="fgss", count_only=True) == c
        ms.append(ctx.stats()["total_ms"])
    st = ctx.stats()
    res = {"long_rs_pairs": int(c), "long_rs_fgss_ms": round(min(ms), 2),
           "long_rs_posting_visits": int(st["posting_visits"]), "long_rs_signature_tests": int(st["signature_tests"]),
           "long_rs_signature_bits": int(st["signature_bits"]), "long_rs_candidates": int(st["candidates"]),
           "long_rs_tile_plan": int(st["tile_plan"])}
    log(f"long R-S cosine 0.8: {c} pairs, fgss {res['long_rs_fgss_ms']} ms, {st['posting_visits']} visits, "
        f"{st['signature_tests']} signature tests ({st['signature_bits']} bits), {st['candidates']} verified")
    assert ctx.join(dS, dR, algo="fgss", tune=H.Tune(flags=H.TUNE_NO_SIGNATURE), count_only=True) == c
    res["long_rs_fgss_no_signature_ms"] = round(ctx.stats()["total_ms"], 1)
    log(f"long R-S without the signature test: {res['long_rs_fgss_no_signature_ms']} ms")
    # the bitmap technique with 1024-bit signatures: every pair of the length
    # windows tested, the first 512 bits per round, survivors merged
    bt = H.Tune(bitmap_bits=1024, flags=H.TUNE_STAGE_TIMES)
    ms = []
    for _ in range(3):
        assert ctx.join(dS, dR, algo="bitmap", tune=bt, count_only=True) == c
        ms.append(ctx.stats()["total_ms"])
    st = ctx.stats()
    res.update({"long_rs_bitmap1024_ms": round(min(ms), 2), "long_rs_bitmap1024_pair_tests": int(st["posting_visits"]),
                "long_rs_bitmap1024_survivors": int(st["candidates"])})
    log(f"long R-S bitmap 1024 bits: {res['long_rs_bitmap1024_ms']} ms, {st['posting_visits']} pair tests, "
        f"{st['candidates']} survivors")
    res.update(bench_coprocess(ctx, dS, dR, "long_rs", 0, c))
    dR.free()
    dS.free()
    ctx.set_threshold(0.8)
    return res


def bench_batched(ctx, H, split=True) -> dict:
    """512 independent Jaccard-0.75 self-joins (the batch workload): one
    persistent launch over all of them vs one hs_join per join."""
    import numpy as np

    from hyset_amd import workloads as W
    t0 = time.perf_counter()
    colls = W.batch()
    log(f"batch: {len(colls)} joins, {sum(c.n for c in colls)} sets, "
        f"{sum(int(c.tokens.shape[0]) for c in colls)} tokens, generated in {time.perf_counter() - t0:.1f}s")
    ctx.set_threshold(0.75)
    singles = [ctx.upload(c) for c in colls]
    b = ctx.upload_batch(colls)
    # plain: one join (index build + launch + count read-back) per collection
    for d in singles[:8]:
        ctx.join(d, count_only=True)
    plain_counts, plain_ms = [], []
    for rep in range(2):
        a = time.perf_counter()
        counts = [ctx.join(d, count_only=True) for d in singles]
        plain_ms.append((time.perf_counter() - a) * 1e3)
        plain_counts = counts
    plain = min(plain_ms)
    total = sum(plain_counts)
    res = {"batch_joins": len(colls), "batch_sets": int(sum(c.n for c in colls)), "batch_pairs": int(total),
           "batch_plain_ms": round(plain, 2)}

    def timed(tune, reps=3):
        ctx.join_batch(b, tune=tune, count_only=True)
        ms = []
        for _ in range(reps):
            a = time.perf_counter()
            c = ctx.join_batch(b, tune=tune, count_only=True)
            ms.append((time.perf_counter() - a) * 1e3)
        return c, min(ms)

    c, ms = timed(None)
    assert c == total, (c, total)
    res["batch_persistent_ms"] = round(ms, 2)
    c, ms_g = timed(H.Tune(flags=H.TUNE_GRAPH))
    assert c == total
    res["batch_persistent_graph_ms"] = round(ms_g, 2)
    tune, _ = ctx.autotune_batch(b, reps=3)
    tune.flags |= H.TUNE_GRAPH  # (the tuned flags kept: lane width, prefix scheme, signature test)
    c, ms_t = timed(tune)
    assert c == total
    res["batch_autotuned_ms"] = round(ms_t, 2)
    res["batch_autotuned_config"] = {"probes_per_tile": tune.probes_per_tile, "tile_tokens": tune.tile_tokens,
                                     "workgroups_per_cu": tune.workgroups_per_cu, "signature_bits": tune.signature_bits,
                                     "narrow": bool(tune.flags & H.TUNE_NARROW),
                                     "index_maxprefix": bool(tune.flags & H.TUNE_INDEX_MAXPREFIX),
                                     "no_signature": bool(tune.flags & H.TUNE_NO_SIGNATURE)}
    best = min(ms, ms_g, ms_t)
    res["batch_speedup_vs_plain"] = round(plain / best, 2)
    # per-join result parity of the batched form against the plain form on a sample of joins
    r, s, ov = ctx.join_batch(b)
    jf = b.join_first
    jr = np.searchsorted(jf, r, side="right") - 1
    ok = bool(np.array_equal(np.bincount(jr, minlength=len(colls)), np.array(plain_counts)))
    for j in (0, 7, 100, 311, 511):
        m = jr == j
        want = H.sort_pairs(*ctx.join(singles[j]))
        got = H.sort_pairs(r[m] - jf[j], s[m] - jf[j], ov[m])
        ok = ok and all(np.array_equal(x, y) for x, y in zip(got, want))
    res["batch_parity_vs_plain"] = ok
    log(f"batch: plain {plain:.1f} ms, persistent {ms:.1f} ms, +graph {ms_g:.1f} ms, autotuned {ms_t:.1f} ms "
        f"({res['batch_autotuned_config']}), speedup {res['batch_speedup_vs_plain']}x, parity={ok}")
    b.free()
    if split:
        res.update(bench_split(ctx, H, singles, np.array(plain_counts, dtype=np.uint64)))
    for d in singles:
        d.free()
    ctx.set_threshold(0.8)
    return res


def bench_split(ctx, H, units, want_counts) -> dict:
    """Learned vs analytic CPU/GPU split on the batch mix (512 Jaccard-0.75
    self-joins).  Both models are trained on timings of a disjoint training
    mix (192 joins drawn like the batch mix, seed 1001), then plan the batch
    mix; every unit's four techniques are timed to score the plans, and each
    plan is executed by hs_split_run (device units on the context stream,
    CPU units on host threads) with its pair counts checked."""
    import numpy as np

    from hyset_amd import split as S
    from hyset_amd import workloads as W
    threads = S.default_threads()
    t0 = time.perf_counter()
    train = [ctx.upload(c) for c in W.batch(k=192, seed=1001)]
    Ftr, Ttr, _ = S.measure_units(ctx, train, 0.75)
    cal = S.calibrate_concurrency(ctx, train, Ttr, threads)
    cpu_scale = cal["cpu"]
    for d in train:
        d.free()
    t_train_mix = time.perf_counter() - t0
    a = time.perf_counter()
    sp = S.Split(ctx).fit(Ftr, Ttr)
    t_fit = time.perf_counter() - a
    F, T, counts = S.measure_units(ctx, units, 0.75)
    assert np.array_equal(counts.astype(np.uint64), want_counts)
    log(f"split: timed {len(train)} training + {len(units)} batch joins x 4 techniques in "
        f"{time.perf_counter() - t0:.1f}s; concurrency scales {cal}; fit {t_fit:.1f}s, "
        f"held-out log-MSE {sp.val}")
    pa, pl = sp.analytic.predict(F), sp.learned.predict(F)
    plans = {"all_fgss": np.full(len(units), H.TECH_FGSS, dtype=np.int32),
             "analytic_argmin": S.argmin_plan(pa), "learned_argmin": S.argmin_plan(pl),
             "analytic_concurrent": S.concurrent_plan(pa, threads, cpu_scale, cal["gpu_loaded"], cal["gpu_alone"]),
             "learned_concurrent": S.concurrent_plan(pl, threads, cpu_scale, cal["gpu_loaded"], cal["gpu_alone"]),
             "measured_best_concurrent": S.concurrent_plan(T, threads, cpu_scale, cal["gpu_loaded"],
                                                           cal["gpu_alone"])}
    summary = S.summarize(plans, T, threads, cpu_scale, cal["gpu_loaded"], cal["gpu_alone"])
    for name, plan in plans.items():
        got, unit_ms, wall = ctx.split_run(units, plan, cpu_threads=threads)
        assert np.array_equal(got, want_counts), name
        summary[name]["run_wall_ms"] = round(wall, 2)
        on_cpu = plan == H.TECH_CPU
        # the two sides of the run: the device queue's time, and the host
        # threads' busy time spread over the threads
        summary[name]["run_device_ms"] = round(float(unit_ms[~on_cpu].sum()), 2)
        summary[name]["run_cpu_ms_per_thread"] = round(float(unit_ms[on_cpu].sum()) / threads, 2)
        log(f"split {name}: run {wall:.1f} ms (device side {summary[name]['run_device_ms']} ms, CPU side "
            f"{summary[name]['run_cpu_ms_per_thread']} ms per thread), modelled sum {summary[name]['sum_ms']} ms "
            f"(regret {summary[name]['regret_ms']}), makespan {summary[name]['makespan_ms']} ms, "
            f"share {summary[name]['share']}")
    judge = sp.judge(F, T)
    log(f"split judge: {judge}")
    return {"split": {"cpu_threads": threads, "concurrency_scales": {k: round(v, 3) for k, v in cal.items()},
                      "training_units": len(train),
                      "units": len(units),
                      "training_mix_s": round(t_train_mix, 1), "fit_s": round(t_fit, 2),
                      "heldout_log_mse": {k: round(v, 4) for k, v in sp.val.items()},
                      "eval_log_mse": {"analytic": round(S.log_error(pa, T), 4),
                                       "learned": round(S.log_error(pl, T), 4)},
                      "uses_learned": sp.use_learned, "plans": summary, "judge": judge}}


def bench_serving(H, slots=4, max_batch_queries=65536, n_requests=2000, offered=(40_000.0, 80_000.0)) -> dict:
    """Serving workload: device-resident index over 2,000,000 sets; requests of
    round(exp(U(0, ln 4096))) query sets at thresholds drawn from {0.6, 0.7,
    0.8, 0.9}.  Saturated throughput (every request queued at once), then
    Poisson arrivals at each fixed offered rate of `offered` requests/s and
    at half the saturated rate; then the saturated run again under a device
    budget of 1.5x the largest request's bytes (requests wait for bytes),
    with every request's pair count checked against the unconstrained run."""
    import random
    import threading

    import numpy as np

    from hyset_amd import workloads as W
    t0 = time.perf_counter()
    idx, pool = W.serving()
    log(f"serving: index {idx.n} sets, query pool {pool.n} sets, generated in {time.perf_counter() - t0:.1f}s")
    rnd = random.Random(2)
    reqs, pos = [], 0
    for _ in range(n_requests):
        n = max(1, int(round(np.exp(rnd.uniform(0, np.log(4096))))))
        if pos + n > pool.n:
            pos = 0
        st = pool.starts[pos]
        en = pool.starts[pos + n - 1] + pool.sizes[pos + n - 1]
        qc = H.SetCollection(pool.tokens[st:en], pool.starts[pos:pos + n] - st, pool.sizes[pos:pos + n], pool.universe)
        reqs.append((qc, rnd.choice([0.6, 0.7, 0.8, 0.9])))
        pos += n
    ctx = H.Context(0)
    d = ctx.upload(idx)
    res = {"serve_index_sets": idx.n, "serve_requests": len(reqs),
           "serve_queries": int(sum(q.n for q, _ in reqs))}
    res.update(serve_slots=slots, serve_max_batch_queries=max_batch_queries)

    def saturated(srv):
        a = time.perf_counter()
        ids = [srv.submit(qc, t) for qc, t in reqs]
        counts = [srv.wait(i)[0] for i in ids]
        return counts, time.perf_counter() - a

    def offered_run(srv, lam, seed):
        prng = random.Random(seed)
        gaps = [prng.expovariate(lam) for _ in reqs]
        lat, done_pairs, ids = [], [0], []
        lock = threading.Lock()

        def waiter():
            k = 0
            while True:
                with lock:
                    if k >= len(ids):
                        if len(ids) == len(reqs):
                            return
                        rid = None
                    else:
                        rid = ids[k]
                if rid is None:
                    time.sleep(0.0002)
                    continue
                c, l = srv.wait(rid)
                lat.append(l)
                done_pairs[0] += c
                k += 1

        s0 = srv.stats()
        th = threading.Thread(target=waiter)
        th.start()
        a = time.perf_counter()
        nxt = a
        for (qc, t), g in zip(reqs, gaps):
            nxt += g
            dt = nxt - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            rid = srv.submit(qc, t)
            with lock:
                ids.append(rid)
        th.join()
        el = time.perf_counter() - a
        s1 = srv.stats()
        la = np.array(lat)
        return {"offered_req_per_s": round(lam, 1), "achieved_req_per_s": round(len(reqs) / el, 1),
                "p50_ms": round(float(np.percentile(la, 50)), 3), "p99_ms": round(float(np.percentile(la, 99)), 3),
                "pairs_per_s": round(done_pairs[0] / el, 1), "batches": s1["batches"] - s0["batches"],
                "max_in_flight": s1["max_in_flight"]}

    with H.Server(d, slots=slots, hbm_budget_bytes=2 << 30, max_batch_queries=max_batch_queries) as srv:
        for qc, t in reqs[:20]:
            srv.wait(srv.submit(qc, t))
        counts, sat_s = saturated(srv)
        pairs = sum(counts)
        sat_rps = len(reqs) / sat_s
        res.update(serve_saturated_req_per_s=round(sat_rps, 1), serve_saturated_ms=round(sat_s * 1e3, 1),
                   serve_pairs=int(pairs), serve_saturated_pairs_per_s=round(pairs / sat_s, 1))
        log(f"serving: saturated {sat_rps:.0f} req/s ({pairs / sat_s / 1e6:.1f}M pairs/s)")
        loads = {}
        for k, lam in enumerate(list(offered) + [0.5 * sat_rps]):
            r = offered_run(srv, lam, 3 + k)
            key = "half_saturated" if k == len(offered) else f"{int(lam / 1000)}k"
            loads[key] = r
            log(f"serving at {lam:.0f} req/s offered: achieved {r['achieved_req_per_s']:.0f}, p50 {r['p50_ms']} ms, "
                f"p99 {r['p99_ms']} ms, {r['batches']} batches, max {r['max_in_flight']} in flight")
        sig_bytes = srv.stats()["signature_bits"] // 8
        st = srv.stats()
    res["serve_offered_loads"] = loads
    # compatibility with earlier lines: the half-saturated run's figures
    h = loads["half_saturated"]
    res.update(serve_offered_req_per_s=h["offered_req_per_s"], serve_achieved_req_per_s=h["achieved_req_per_s"],
               serve_p50_ms=h["p50_ms"], serve_p99_ms=h["p99_ms"], serve_pairs_per_s=h["pairs_per_s"],
               serve_batches=st["batches"], serve_max_in_flight=st["max_in_flight"], serve_deferred=st["deferred"])
    # the budget run: room for 1.5 of the largest request at a time
    need = max(4 * int(qc.tokens.shape[0]) + (109 + sig_bytes) * qc.n + 4096 for qc, _ in reqs)
    budget = need * 3 // 2
    with H.Server(d, slots=slots, hbm_budget_bytes=budget, max_batch_queries=max_batch_queries) as srv:
        bcounts, b_s = saturated(srv)
        bst = srv.stats()
    same = bcounts == counts
    res["serve_budget"] = {"hbm_budget_bytes": int(budget), "largest_request_bytes": int(need),
                           "req_per_s": round(len(reqs) / b_s, 1), "deferred": int(bst["deferred"]),
                           "rejected": int(bst["rejected"]), "max_bytes_in_flight": int(bst["max_bytes_in_flight"]),
                           "answers_unchanged": bool(same)}
    log(f"serving under a {budget / 1e6:.1f} MB budget: {len(reqs) / b_s:.0f} req/s, {bst['deferred']} deferred, "
        f"max {bst['max_bytes_in_flight'] / 1e6:.1f} MB in flight, answers unchanged: {same}")
    assert same and bst["max_bytes_in_flight"] <= budget
    ctx.close()
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--quick", action="store_true", help="headline only: no side measurements, no CPU baseline")
    ap.add_argument("--threshold", type=float, default=0.8)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-join measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-split", action="store_true", help="skip the learned/analytic split measurement")
    ap.add_argument("--no-long", action="store_true", help="skip the long-set R-S measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        return 2

    import numpy as np

    import hyset_amd as H

    t = args.threshold
    log("generating 1M-set collection")
    g0 = time.perf_counter()
    coll = H.generate_self(H.gen_params(1_000_000, median=10, sigma=0.6, lo=2, hi=200, universe=1_000_000,
                                        zipf=1.1, seed=1))
    log(f"generated n={coll.n} tokens={coll.tokens.shape[0]} in {time.perf_counter() - g0:.1f}s")
    ctx = H.Context(0)
    d = ctx.upload(coll)
    ctx.set_threshold(t)
    copy_gbps = ctx.copy_bandwidth(1 << 30, 10)
    log(f"device copy kernel: {copy_gbps:.0f} GB/s")

    for _ in range(args.warmup):
        ctx.join(d, algo="fgss", count_only=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pairs = ctx.join(d, algo="fgss", count_only=True)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    # the device time of the same joins (read after each join, outside the timed loop)
    steps = []
    for _ in range(args.steps):
        assert ctx.join(d, algo="fgss", count_only=True) == pairs
        steps.append(ctx.stats())
    st = steps[-1]
    dev_ms = statistics.median(s["total_ms"] for s in steps)
    # The same steps again with the stage events (HS_TUNE_STAGE_TIMES): the
    # index build's end and the join kernels' start and end (the token-major
    # filter's start to the verification's end), taken by the dispatches
    # themselves on the join's stream.  They cost the device a
    # few microseconds per join, so the throughput above runs without them.
    stune = H.Tune(flags=H.TUNE_STAGE_TIMES)
    ctx.join(d, algo="fgss", tune=stune, count_only=True)
    a = time.perf_counter()
    for _ in range(args.steps):
        assert ctx.join(d, algo="fgss", tune=stune, count_only=True) == pairs
    ms_per_step_stages = (time.perf_counter() - a) * 1e3 / args.steps
    ssteps = []
    for _ in range(args.steps):
        assert ctx.join(d, algo="fgss", tune=stune, count_only=True) == pairs
        ssteps.append(ctx.stats())
    kern_ms = statistics.median(s["kernel_ms"] for s in ssteps)
    idx_ms = statistics.median(s["index_ms"] for s in ssteps)
    log(f"fgss t={t}: {pairs} pairs, {ms_per_step:.3f} ms/step wall, device {dev_ms:.3f} ms; with stage events "
        f"{ms_per_step_stages:.3f} ms/step (index {idx_ms:.3f}, join kernels {kern_ms:.3f})")

    # the same join replayed from a HIP graph (HS_TUNE_GRAPH): every kernel
    # runs each step; the host enqueues one graph instead of the sequence
    gtune = H.Tune(flags=H.TUNE_GRAPH)
    for _ in range(max(args.warmup, 2)):
        assert ctx.join(d, algo="fgss", tune=gtune, count_only=True) == pairs
    a = time.perf_counter()
    for _ in range(args.steps):
        assert ctx.join(d, algo="fgss", tune=gtune, count_only=True) == pairs
    ms_per_step_graph = (time.perf_counter() - a) * 1e3 / args.steps
    gsteps = []
    for _ in range(args.steps):
        assert ctx.join(d, algo="fgss", tune=gtune, count_only=True) == pairs
        gsteps.append(ctx.stats())
    replayed = all(g["graph"] == 2 for g in gsteps)
    dev_ms_graph = statistics.median(g["total_ms"] for g in gsteps)
    log(f"fgss t={t} graph replay: {ms_per_step_graph:.3f} ms/step wall, device {dev_ms_graph:.3f} ms "
        f"(replayed every step: {replayed})")

    # with every result triple copied into host memory inside the timed
    # region (the pairs' device-to-host copy rides with the counters' copy
    # when the result fits the previous join's size)
    for _ in range(2):
        ctx.join(d, algo="fgss")
    n_d2h = max(5, args.steps // 4)
    a = time.perf_counter()
    for _ in range(n_d2h):
        r, s, ov = ctx.join(d, algo="fgss")
    ms_with_d2h = (time.perf_counter() - a) * 1e3 / n_d2h
    assert r.shape[0] == pairs
    dev_ms_d2h = ctx.stats()["total_ms"]
    log(f"fgss t={t} with the pairs in host memory: {ms_with_d2h:.3f} ms/step wall, device {dev_ms_d2h:.3f} ms")

    sst = ssteps[-1]  # stage-timed: also counts |r| over the probes with candidates
    tok = sst["tile_plan"] == H.PLAN_TOKEN
    sb = stage_bytes(sst, coll.n, coll.universe, int(coll.tokens.shape[0]))
    alg = sb["candidates"] + sb["verify"]
    achieved = alg / (kern_ms * 1e-3) / 1e9
    e2e = sum(sb.values())
    e2e_gbps = e2e / (dev_ms * 1e-3) / 1e9
    design = token_join_design_bytes(sst) if tok else None
    traffic = None
    tfile = ROOT / "profiles" / "tok_join_1M_jaccard0.8_traffic.json"
    if tok and tfile.exists():
        traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
    log(f"stage bytes {sb}; join kernels {achieved:.0f} GB/s, end to end {e2e_gbps:.0f} GB/s")

    extra = {"ms_per_step_graph": round(ms_per_step_graph, 4), "join_ms_device_graph": round(dev_ms_graph, 4),
             "graph_replayed_every_step": replayed,
             "join_ms_device": round(dev_ms, 4), "index_ms": round(idx_ms, 4), "join_kernels_ms": round(kern_ms, 4),
             "ms_per_step_stage_events": round(ms_per_step_stages, 4),
             "value_with_d2h": round(pairs / (ms_with_d2h * 1e-3), 1), "ms_per_step_with_d2h": round(ms_with_d2h, 4),
             "join_ms_device_with_d2h": round(dev_ms_d2h, 4), "pairs": int(pairs),
             "end_to_end": {"algorithmic_bytes": int(e2e), "stage_bytes": {k: int(v) for k, v in sb.items()},
                            "ms": round(dev_ms, 4), "achieved_gbps": round(e2e_gbps, 1),
                            "frac": round(e2e_gbps / HBM_PEAK_GBPS, 4)},
             "candidates": int(sst["candidates"]), "posting_visits": int(sst["posting_visits"]),
             "index_entries": int(sst["index_entries"]), "probe_prefix_tokens": int(sst["probe_prefix_tokens"]),
             "candidate_tokens": int(sst["candidate_tokens"]),
             "candidate_probe_tokens": int(sst["candidate_probe_tokens"]),
             "signature_tests": int(sst["signature_tests"]), "signature_bits": int(sst["signature_bits"]),
             "tile_plan": int(sst["tile_plan"]),
             "hbm_copy_kernel_gbps": round(copy_gbps, 1)}

    cpu = None
    if not args.quick:
        import oracle as O
        # parity of the measured result against the oracle (cpu_single restated)
        r, s, ov = H.sort_pairs(r, s, ov)
        want = O.self_join(coll, t)  # warm-up
        extra["parity_vs_oracle"] = bool(np.array_equal(r, want.r) and np.array_equal(s, want.s)
                                         and np.array_equal(ov, want.overlap))
        # the reference's intervals on a steady clock: "join time" without
        # the index build, "total time without I/O" with it; median of 5
        # after the warm-up
        runs = []
        for _ in range(5):
            w = O.self_join(coll, t)
            assert w.count == want.count
            runs.append((w.index_ms, w.join_ms))
        join_ms = statistics.median(j for _, j in runs)
        total_ms = statistics.median(i + j for i, j in runs)
        ncores = max(1, min(16, len(os.sched_getaffinity(0))))  # the box's CPU share for one GPU
        model = cpu_model()
        cpu = {"value": round(want.count / (total_ms * 1e-3), 1), "unit": "pairs/s", "cores": 1, "kind": "port",
               "join_ms": round(join_ms, 2), "total_ms": round(total_ms, 2), "cpu_model": model,
               "nproc": os.cpu_count(), "usable_cpus": ncores,
               "sample": f"full workload: cpu_single (AllPairs restated in C), 1M sets, t={t}; steady clock, median "
                         f"of 5 after 1 warm-up: join time {join_ms:.1f} ms, total without I/O (index build + "
                         f"join) {total_ms:.1f} ms; value = pairs / total; {model}, nproc {os.cpu_count()}"}
        O.cpu_multi(coll, t, ncores)  # warm-up
        mruns = []
        for _ in range(5):
            a = time.perf_counter()
            w = O.cpu_multi(coll, t, ncores)
            mruns.append((w.index_ms, w.join_ms, (time.perf_counter() - a) * 1e3))
        extra["cpu_multi"] = {"threads": ncores, "join_ms": round(statistics.median(m[1] for m in mruns), 2),
                              "index_ms_max_worker": round(statistics.median(m[0] for m in mruns), 2),
                              "wall_ms": round(statistics.median(m[2] for m in mruns), 2),
                              "pairs_per_s": round(want.count / (statistics.median(m[2] for m in mruns) * 1e-3), 1)}
        extra["cpu_single"] = {"join_ms": cpu["join_ms"], "total_ms": cpu["total_ms"]}
        log(f"oracle: {want.count} pairs; cpu_single join {join_ms:.1f} ms, total {total_ms:.1f} ms (1 core); "
            f"cpu_multi {extra['cpu_multi']}; parity={extra['parity_vs_oracle']}")

        # the other thresholds and techniques of the configuration
        for tt in (0.6, 0.7, 0.9):
            ctx.set_threshold(tt)
            for plan, flag in (("", 0), ("_static_tiles", H.TUNE_STATIC_TILES), ("_tile_queue", H.TUNE_TILE_QUEUE),
                               ("_balanced_tiles", H.TUNE_BALANCED_TILES)):
                tu = H.Tune(flags=flag)
                c0 = ctx.join(d, algo="fgss", tune=tu, count_only=True)
                ms = []
                for _ in range(3):
                    c = ctx.join(d, algo="fgss", tune=tu, count_only=True)
                    assert c == c0
                    ms.append(ctx.stats()["total_ms"])
                extra[f"fgss_t{tt}{plan}_ms"] = round(statistics.median(ms), 3)
                if not plan:
                    extra[f"fgss_t{tt}_tile_plan"] = int(ctx.stats()["tile_plan"])
                log(f"fgss t={tt}{plan}: {c} pairs {statistics.median(ms):.3f} ms (plan {ctx.stats()['tile_plan']})")
            extra[f"fgss_t{tt}_pairs"] = c
            if tt == 0.6:
                # the result sink on a large result: the pairs leave the device in
                # pinned chunks, into host arrays or streamed into a text file
                ctx.join(d, algo="fgss", count_only=True)
                cp, tx = [], []
                with tempfile.TemporaryDirectory() as td:
                    tf = Path(td) / "pairs.txt"
                    for _ in range(3):
                        a = time.perf_counter()
                        ctx.last_result(c)
                        cp.append((time.perf_counter() - a) * 1e3)
                        a = time.perf_counter()
                        ctx.write_last_result_text(tf)
                        tx.append((time.perf_counter() - a) * 1e3)
                    extra["sink_t0.6_text_bytes"] = tf.stat().st_size
                extra["sink_t0.6_copy_ms"] = round(statistics.median(cp), 3)
                extra["sink_t0.6_copy_gbps"] = round(12 * c / (statistics.median(cp) * 1e-3) / 1e9, 2)
                extra["sink_t0.6_text_ms"] = round(statistics.median(tx), 3)
                log(f"result sink t=0.6: {c} pairs to host arrays {extra['sink_t0.6_copy_ms']} ms "
                    f"({extra['sink_t0.6_copy_gbps']} GB/s), to text {extra['sink_t0.6_text_ms']} ms")
        ctx.set_threshold(t)
        # the other tile plans: the upload-time tiles on a persistent grid,
        # and work-balanced tiles planned on the device (item-slice pass)
        # (join times without stage events, kernel times with them)
        def plan_times(flags):
            pt, st_ = H.Tune(flags=flags), H.Tune(flags=flags | H.TUNE_STAGE_TIMES)
            ctx.join(d, algo="fgss", tune=pt, count_only=True)
            ks, ds = [], []
            for _ in range(5):
                assert ctx.join(d, algo="fgss", tune=pt, count_only=True) == pairs
                ds.append(ctx.stats()["total_ms"])
                assert ctx.join(d, algo="fgss", tune=st_, count_only=True) == pairs
                ks.append(ctx.stats()["kernel_ms"])
            return round(statistics.median(ks), 4), round(statistics.median(ds), 4), ctx.stats()

        for name, flag in (("item_major", H.TUNE_ITEM_MAJOR), ("static_tiles", H.TUNE_STATIC_TILES),
                           ("tile_queue", H.TUNE_TILE_QUEUE), ("balanced_tiles", H.TUNE_BALANCED_TILES)):
            extra[f"k_fgss_{name}_ms"], extra[f"join_ms_device_{name}"], _ = plan_times(flag)
            log(f"{name}: k_fgss {extra[f'k_fgss_{name}_ms']} ms, join {extra[f'join_ms_device_{name}']} ms "
                f"(token-major: join kernels {kern_ms:.4f} ms, join {dev_ms:.4f} ms)")
        # the same join without the signature test (every positional survivor merged)
        extra["k_fgss_no_signature_ms"], extra["join_ms_device_no_signature"], sn = plan_times(H.TUNE_NO_SIGNATURE)
        extra["candidates_no_signature"] = int(sn["candidates"])
        log(f"k_fgss without the signature test: {extra['k_fgss_no_signature_ms']} ms, "
            f"{extra['candidates_no_signature']} candidates verified")
        # the bitmap filter is bound by integer VALU issue, not HBM: a pair
        # test is an XOR and a popcount-with-accumulate per 32-bit word of
        # the signature's first half (all of it at 64 bits) and one compare
        # -- 2 w + 1 VALU instructions, w = the half's words -- against the
        # measured rate of XOR-popcount chains (2 VALU instructions a step)
        # on this device; rounds where a lane passes the half also test the
        # rest (not counted)
        popc_gops = ctx.popcount_rate()
        extra["popcount_peak_gops"] = round(popc_gops, 1)
        for bits in (64, 128, 256):
            tune = H.Tune(bitmap_bits=bits, flags=H.TUNE_STAGE_TIMES)
            c = ctx.join(d, algo="bitmap", tune=tune, count_only=True)
            sb = ctx.stats()
            extra[f"bitmap{bits}_ms"] = round(sb["total_ms"], 3)
            extra[f"bitmap{bits}_kernel_ms"] = round(sb["kernel_ms"], 3)
            extra[f"bitmap{bits}_pair_tests"] = int(sb["posting_visits"])
            extra[f"bitmap{bits}_survivors"] = int(sb["candidates"])
            half_words = (bits if bits == 64 else bits // 2) // 32
            per_test = 2 * half_words + 1
            ops = sb["posting_visits"] * per_test
            gops = ops / (sb["kernel_ms"] * 1e-3) / 1e9
            extra[f"bitmap{bits}_valu"] = {"valu_per_test": per_test, "achieved_gops": round(gops, 1),
                                           "peak_gops": round(popc_gops, 1), "frac": round(gops / popc_gops, 3),
                                           "tests_per_s": round(sb["posting_visits"] / (sb["kernel_ms"] * 1e-3), 1)}
            assert c == pairs
            log(f"bitmap{bits}: {sb['total_ms']:.2f} ms, {sb['posting_visits']} tests, {sb['candidates']} survivors, "
                f"{gops:.0f} Gop/s of {popc_gops:.0f} ({gops / popc_gops:.2f})")
        a = time.perf_counter()
        c = ctx.join(d, algo="hybrid", tune=H.Tune(host_threads=min(16, ncores)), count_only=True)
        extra["hybrid_ms_wall"] = round((time.perf_counter() - a) * 1e3, 2)
        assert c == pairs
        log(f"hybrid ({min(16, ncores)} host threads): {extra['hybrid_ms_wall']} ms")
        extra.update(bench_coprocess(ctx, d, d, "t0.8", 0, pairs))
        extra.update(bench_policies(ctx, d, pairs, max(2, ncores - 1)))
        extra.update(bench_partition_pairs(ctx, H, d, coll, t, pairs, max(2, ncores - 1)))
        ctx.set_threshold(0.6)
        extra.update(bench_coprocess(ctx, d, d, "t0.6", 0, extra["fgss_t0.6_pairs"]))
        ctx.set_threshold(t)

    if not args.quick and not args.no_long:
        extra.update(bench_long_rs(ctx, H))
    if not args.quick and not args.no_batch:
        extra.update(bench_batched(ctx, H, split=not args.no_split))
    if not args.quick and not args.no_serve:
        extra.update(bench_serving(H))

    out = {
        "metric": "pairs/s (1M-set Jaccard>=0.8 self-join, fgss)",
        "value": round(pairs / (ms_per_step * 1e-3), 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (Zipf/lognormal generator, seed 1)",
        "config": {"workload": "self-join, 1,000,000 sets, lognormal sizes median 10 sigma 0.6 in [2,200], "
                               "Zipf(1.1) over 1M tokens, Jaccard 0.8 (reference double bounds), fgss",
                   "sets": coll.n, "tokens": int(coll.tokens.shape[0]), "threshold": t, "algorithm": "fgss"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "kernel": "token join (k_tok_filter, k_tok_heavy, k_tok_verify)" if tok else "k_fgss",
                     "ms": round(kern_ms, 4), "algorithmic_bytes": int(alg),
                     "design_bytes": int(design) if design is not None else None},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
