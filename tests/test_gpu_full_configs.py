"""Oracle parity at the benchmark configurations' own sizes.

* the 1M-set self-join (sizes lognormal median 10, sigma 0.6 in [2, 200],
  Zipf(1.1) over 1M tokens, seed 1): fgss at Jaccard 0.6, the bitmap
  technique at 64 / 128 / 256 bits and the hybrid technique at 0.8, full
  sorted (r, s, overlap) triples against the oracle's cpu_multi;
* the 512-join batch (Jaccard 0.75 self-joins of 1,000-200,000 sets each,
  20.0M sets): every join's triples of one persistent launch against the
  oracle's join of that collection alone;
* the 2M-set serving index (median 12, Zipf(1.1) over 2M tokens): 64
  requests of 1-4,096 query sets at the four thresholds, each answered alone
  and in a mix, against the oracle's R-S join, also under a device budget
  that makes requests wait;
* Jaccard with exact-rational bounds on the hand-made fixture whose pairs sit
  at similarity exactly 0.8 and 0.9 (the sums where the reference's double
  bound rounds up): every technique keeps exactly the pairs the oracle keeps.
"""
import concurrent.futures
import os
import random
from pathlib import Path

import numpy as np
import pytest

import hyset_amd as H
import oracle as O
from hyset_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _check(got, want, what):
    r, s, ov = H.sort_pairs(*got)
    assert r.shape[0] == want.count, f"{what}: {r.shape[0]} pairs, oracle {want.count}"
    assert np.array_equal(r, want.r), what
    assert np.array_equal(s, want.s), what
    assert np.array_equal(ov, want.overlap), what


@pytest.fixture(scope="module")
def headline(ctx):
    c = W.headline()
    d = ctx.upload(c)
    yield c, d
    d.free()


def test_headline_fgss_t06(ctx, headline):
    """The configuration's loosest threshold (1.42M pairs, balanced tiles)."""
    c, d = headline
    ctx.set_threshold(0.6)
    want = O.cpu_multi(c, 0.6, THREADS)
    assert want.count > 1_000_000
    _check(ctx.join(d, algo="fgss"), want, "fgss 1M t=0.6")
    st = ctx.stats()
    assert st["tile_plan"] == H.PLAN_TOKEN
    assert st["posting_visits"] == want.posting_visits
    sizes, counts = np.unique(c.sizes, return_counts=True)
    assert st["index_entries"] == sum(int(k) * O.bounds("jaccard", 0.6, int(l))[4] for l, k in zip(sizes, counts))
    # the item-major plans at full size (the automatic one is balanced here)
    _check(ctx.join(d, algo="fgss", tune=H.Tune(flags=H.TUNE_ITEM_MAJOR)), want, "fgss 1M t=0.6 item-major")
    assert ctx.stats()["tile_plan"] == H.PLAN_BALANCED
    for flags in (H.TUNE_STATIC_TILES, H.TUNE_TILE_QUEUE):
        _check(ctx.join(d, algo="fgss", tune=H.Tune(flags=flags)), want, f"fgss 1M t=0.6 flags={flags}")


@pytest.mark.parametrize("bits", [64, 128, 256])
def test_headline_bitmap(ctx, headline, bits):
    """The bitmap technique over all 1.1e11 length-window pairs of the 1M
    sets at Jaccard 0.8: full triples, and the filter counts every pair of
    the windows."""
    c, d = headline
    ctx.set_threshold(0.8)
    want = O.self_join(c, 0.8)
    _check(ctx.join(d, algo="bitmap", tune=H.Tune(bitmap_bits=bits)), want, f"bitmap{bits} 1M")
    st = ctx.stats()
    # pairs tested: every (r, s < r) inside r's length window
    sizes = np.asarray(c.sizes, dtype=np.int64)
    lo = np.array([np.searchsorted(sizes, O.bounds("jaccard", 0.8, int(l))[0], side="left") for l in range(201)])
    hi = np.array([np.searchsorted(sizes, O.bounds("jaccard", 0.8, int(l))[1], side="right") for l in range(201)])
    ids = np.arange(c.n, dtype=np.int64)
    win = np.maximum(0, np.minimum(hi[sizes], ids) - lo[sizes])
    assert st["posting_visits"] == int(win.sum())
    assert want.count <= st["candidates"] <= st["posting_visits"]


def test_headline_hybrid(ctx, headline):
    c, d = headline
    ctx.set_threshold(0.8)
    want = O.self_join(c, 0.8)
    _check(ctx.join(d, algo="hybrid", tune=H.Tune(host_threads=THREADS)), want, "hybrid 1M")


def test_batch_512_joins(ctx):
    """One persistent launch over the 512 joins of the batch workload gives,
    join by join, the oracle's triples (ids shifted by each join's first set
    id); so does the autotuned configuration replayed from its graph."""
    colls = W.batch(threads=THREADS)
    assert len(colls) == 512 and sum(c.n for c in colls) > 19_000_000
    with concurrent.futures.ThreadPoolExecutor(THREADS) as ex:
        wants = list(ex.map(lambda c: O.self_join(c, 0.75), colls))
    b = ctx.upload_batch(colls)
    try:
        jf = b.join_first
        want = tuple(np.concatenate([getattr(w, k).astype(np.int64) + (jf[j] if k != "overlap" else 0)
                                     for j, w in enumerate(wants)]).astype(np.uint32) for k in ("r", "s", "overlap"))
        assert want[0].shape[0] > 400_000
        ctx.set_threshold(0.75)

        def same(got, what):
            r, s, ov = H.sort_pairs(*got)
            assert r.shape[0] == want[0].shape[0], f"{what}: {r.shape[0]} pairs, oracle {want[0].shape[0]}"
            assert np.array_equal(r, want[0]) and np.array_equal(s, want[1]) and np.array_equal(ov, want[2]), what

        same(ctx.join_batch(b), "batch")
        tune, _ = ctx.autotune_batch(b, reps=1)
        tune.flags |= H.TUNE_GRAPH
        for rep in range(2):
            same(ctx.join_batch(b, tune=tune), f"autotuned graph rep {rep}")
    finally:
        b.free()


# ------------------------------------------------------------------ serving

N_REQUESTS = 64
THRESHOLDS = (0.6, 0.7, 0.8, 0.9)


@pytest.fixture(scope="module")
def serving(ctx):
    idx, pool = W.serving()
    d = ctx.upload(idx)
    rnd = random.Random(5)
    reqs = []
    for k in range(N_REQUESTS):
        # sizes as the serving workload draws them, with both ends present
        n = 1 if k == 0 else 4096 if k == 1 else max(1, int(round(np.exp(rnd.uniform(0, np.log(4096))))))
        pos = rnd.randrange(0, pool.n - n)
        st, en = int(pool.starts[pos]), int(pool.starts[pos + n - 1] + pool.sizes[pos + n - 1])
        qc = H.SetCollection(pool.tokens[st:en], pool.starts[pos:pos + n] - st, pool.sizes[pos:pos + n],
                             pool.universe)
        reqs.append((qc, THRESHOLDS[k % 4]))
    # the oracle: per threshold, one R-S join of all its requests' queries
    want = [None] * len(reqs)
    for t in THRESHOLDS:
        ks = [k for k, (_, tt) in enumerate(reqs) if tt == t]
        toks = np.concatenate([reqs[k][0].tokens for k in ks])
        sizes = np.concatenate([reqs[k][0].sizes for k in ks])
        starts = np.zeros_like(sizes)
        starts[1:] = np.cumsum(sizes[:-1], dtype=np.uint64).astype(np.uint32)
        res = O.rs_join(H.SetCollection(toks, starts, sizes, pool.universe), idx, t)
        first = np.concatenate([[0], np.cumsum([reqs[k][0].n for k in ks])])
        which = np.searchsorted(first, res.r, side="right") - 1
        for i, k in enumerate(ks):
            m = which == i
            want[k] = list(zip((res.r[m] - first[i]).tolist(), res.s[m].tolist(), res.overlap[m].tolist()))
    yield idx, d, reqs, want
    d.free()


def _triples(r, s, ov):
    r, s, ov = H.sort_pairs(r, s, ov)
    return list(zip(r.tolist(), s.tolist(), ov.tolist()))


def test_serving_2m_alone_and_mix(serving):
    idx, d, reqs, want = serving
    assert sum(len(w) for w in want) > 2_000
    assert {t for _, t in reqs} == set(THRESHOLDS)
    with H.Server(d, slots=4) as srv:
        alone = [_triples(*srv.wait_pairs(srv.submit(qc, t))[:3]) for qc, t in reqs]
        ids = [srv.submit(qc, t) for qc, t in reqs]
        mix = [_triples(*srv.wait_pairs(i)[:3]) for i in ids]
        st = srv.stats()
    for k, (a, m, w) in enumerate(zip(alone, mix, want)):
        assert a == w, f"request {k} alone ({reqs[k][0].n} queries, t={reqs[k][1]})"
        assert m == w, f"request {k} in the mix"
    assert st["batches"] < 2 * N_REQUESTS


def test_serving_2m_budget_defers(serving):
    """A device budget of 1.5 times the largest request: requests wait for
    the bytes held by batches in flight, the bytes in flight never exceed
    the budget, and every answer is unchanged."""
    idx, d, reqs, want = serving
    with H.Server(d, slots=1) as probe_srv:
        sig_bytes = probe_srv.stats()["signature_bits"] // 8
    need = max(4 * int(qc.tokens.shape[0]) + (109 + sig_bytes) * qc.n + 4096 for qc, _ in reqs)
    budget = need * 3 // 2
    # the 4096-query requests first, so the first batches are long and the
    # next requests find their bytes taken
    order = sorted(range(len(reqs)), key=lambda k: -reqs[k][0].n)
    with H.Server(d, slots=4, hbm_budget_bytes=budget) as srv:
        ids = {k: srv.submit(reqs[k][0], reqs[k][1]) for k in order}
        got = {k: _triples(*srv.wait_pairs(ids[k])[:3]) for k in order}
        st = srv.stats()
    for k in order:
        assert got[k] == want[k], f"request {k} under the budget"
    assert st["max_bytes_in_flight"] <= budget
    assert st["deferred"] > 0 and st["rejected"] == 0


# ------------------------------------------------------------- exact bounds

@pytest.mark.parametrize("algo", ["fgss", "bitmap", "hybrid"])
def test_exact_bounds_special_fixture(ctx, algo):
    """HS_BOUNDS_EXACT keeps the pairs whose Jaccard similarity equals the
    threshold at the sums where the reference's double quotient rounds up
    (31 + 32 = 63 at 0.8, 19 + 19 = 38 at 0.9); the reference bounds drop
    exactly those.  Both modes against the oracle and its brute force."""
    c = H.read_collection_text(GOLD / "special.txt")
    d = ctx.upload(c)
    try:
        for t in (0.8, 0.9):
            ref = O.self_join(c, t, mode=O.BOUNDS_REFERENCE)
            exact = O.self_join(c, t, mode=O.BOUNDS_EXACT)
            brute = O.bruteforce_self(c, t, mode=O.BOUNDS_EXACT)
            assert np.array_equal(exact.r, brute.r) and np.array_equal(exact.s, brute.s)
            extra = set(zip(exact.r.tolist(), exact.s.tolist())) - set(zip(ref.r.tolist(), ref.s.tolist()))
            assert extra, t  # the fixture holds pairs at similarity exactly t
            for r_, s_ in extra:
                ov = O.intersect(c.set(r_), c.set(s_))
                lr, ls = int(c.sizes[r_]), int(c.sizes[s_])
                assert ov * (1 + t) == pytest.approx(t * (lr + ls))  # J = ov / (lr + ls - ov) = t
            ctx.set_threshold(t, "jaccard", H.BOUNDS_EXACT)
            _check(ctx.join(d, algo=algo), exact, f"{algo} exact t={t}")
            ctx.set_threshold(t, "jaccard", H.BOUNDS_REFERENCE)
            _check(ctx.join(d, algo=algo), ref, f"{algo} reference t={t}")
    finally:
        d.free()
