"""GPU parity: every technique returns exactly the oracle's sorted (r, s,
overlap) triples (the reference's cpu_single result plus exact overlaps)."""
import numpy as np
import pytest

import hyset_amd as H
import oracle as O

pytestmark = pytest.mark.gpu


def _check(got, want, what):
    r, s, ov = H.sort_pairs(*got)
    assert r.shape[0] == want.count, f"{what}: {r.shape[0]} pairs, oracle {want.count}"
    assert np.array_equal(r, want.r), what
    assert np.array_equal(s, want.s), what
    assert np.array_equal(ov, want.overlap), what


@pytest.fixture(scope="module")
def small(ctx):
    c = H.generate_self(H.gen_params(3000, size_dist=H.SIZE_UNIFORM, lo=2, hi=40, universe=800, zipf=1.0,
                                     seed=5))
    return c, ctx.upload(c)


@pytest.mark.parametrize("t", [0.6, 0.7, 0.75, 0.8, 0.9])
@pytest.mark.parametrize("sim", ["jaccard", "cosine", "dice"])
def test_fgss_small(ctx, small, t, sim):
    c, d = small
    ctx.set_threshold(t, sim)
    want = O.self_join(c, t, sim)
    _check(ctx.join(d, algo="fgss"), want, f"fgss {sim} {t}")


@pytest.mark.parametrize("tune", ["default", "item", "static", "queue", "balanced", "ppt64", "ppt1", "wpc1"])
@pytest.mark.parametrize("t", [0.6, 0.8])
def test_fgss_tile_plans(ctx, small, t, tune):
    """Every plan gives the oracle's result: the automatic choice
    (token-major for a self-join of one range), item-major with its
    automatic tiles, one workgroup per upload-time tile, the same tiles on a
    persistent grid (also a small one), the work-balanced plan, fixed-size
    tiles; whole range and sub-ranges."""
    c, d = small
    ctx.set_threshold(t)
    want = O.self_join(c, t)
    tu = {"default": H.Tune(), "item": H.Tune(flags=H.TUNE_ITEM_MAJOR), "static": H.Tune(flags=H.TUNE_STATIC_TILES),
          "queue": H.Tune(flags=H.TUNE_TILE_QUEUE), "balanced": H.Tune(flags=H.TUNE_BALANCED_TILES),
          "ppt64": H.Tune(probes_per_tile=64), "ppt1": H.Tune(probes_per_tile=1),
          "wpc1": H.Tune(flags=H.TUNE_TILE_QUEUE, workgroups_per_cu=1)}[tune]
    _check(ctx.join(d, algo="fgss", tune=tu), want, f"fgss {tune} {t}")
    plan = {"default": H.PLAN_TOKEN, "item": H.PLAN_UPLOAD, "static": H.PLAN_UPLOAD, "queue": H.PLAN_QUEUE,
            "balanced": H.PLAN_BALANCED, "ppt64": H.PLAN_FIXED, "ppt1": H.PLAN_FIXED, "wpc1": H.PLAN_QUEUE}[tune]
    st = ctx.stats()
    assert st["tile_plan"] == plan  # 3000 sets: too few items for the automatic balanced plan
    # every plan counts the same work: postings inside the windows, the index's entries
    assert st["posting_visits"] == want.posting_visits and st["index_entries"] == want.index_entries
    for rng in ((0, 0), (1, 1), (2999, 2999), (100, 2500)):
        want = O.self_join_blocks(c, t, rng, rng)
        _check(ctx.join(d, algo="fgss", tune=tu, indexed_range=rng, probe_range=rng), want, f"fgss {tune} {rng}")


@pytest.mark.parametrize("bits", [64, 128, 256, 512, 1024, "none"])
@pytest.mark.parametrize("t,sim", [(0.6, "jaccard"), (0.8, "jaccard"), (0.9, "jaccard"), (0.7, "cosine"),
                                   (0.8, "dice")])
def test_fgss_signature_widths(ctx, small, t, sim, bits):
    """The signature test between the positional filter and verification
    only drops visits that cannot pass: every width, and no test at all,
    give the oracle's result; the test runs and prunes."""
    c, d = small
    ctx.set_threshold(t, sim)
    want = O.self_join(c, t, sim)
    tu = H.Tune(flags=H.TUNE_NO_SIGNATURE) if bits == "none" else H.Tune(signature_bits=bits)
    _check(ctx.join(d, algo="fgss", tune=tu), want, f"fgss signature {bits} {sim} {t}")
    st = ctx.stats()
    if bits == "none":
        assert st["signature_bits"] == 0 and st["signature_tests"] == 0
    else:
        assert st["signature_bits"] == bits
        assert st["signature_tests"] >= st["candidates"]
    rng = (100, 2500)
    _check(ctx.join(d, algo="fgss", tune=tu, indexed_range=rng, probe_range=rng),
           O.self_join_blocks(c, t, rng, rng, sim), f"fgss signature {bits} block {rng}")


@pytest.mark.parametrize("n", [3000, 10000])
def test_index_long_lists(ctx, n):
    """Posting lists longer than a workgroup sorts in LDS (2048 keys): every
    set shares tokens 1 and 2 (not frequency-ranked on purpose), so lists 1
    and 2 hold every set and are sorted in runs and merged.  Results, index
    size and posting visits equal the oracle's."""
    rng = np.random.default_rng(n)
    sets = []
    for i in range(n):
        k = 3 + i * 6 // n  # sizes 3..8, ascending
        rest = np.sort(rng.choice(np.arange(3, 400), size=k - 2, replace=False))
        sets.append([1, 2] + rest.tolist())
    c = H.SetCollection.from_sets(sets)
    d = ctx.upload(c)
    try:
        for t in (0.5, 0.6, 0.8):
            ctx.set_threshold(t)
            want = O.self_join(c, t)
            _check(ctx.join(d, algo="fgss"), want, f"long lists n={n} t={t}")
            st = ctx.stats()
            assert st["index_entries"] == want.index_entries
            assert st["posting_visits"] == want.posting_visits
    finally:
        d.free()


def test_fgss_signature_prunes_candidates(ctx):
    """On the 1M-set workload's distribution (50K sets) the 64-bit signature
    test leaves far fewer merges than the positional filter alone, and the
    same pairs."""
    c = H.generate_self(H.gen_params(50_000, median=10, sigma=0.6, lo=2, hi=200, universe=50_000, zipf=1.1,
                                     seed=3))
    d = ctx.upload(c)
    try:
        ctx.set_threshold(0.8)
        a = H.sort_pairs(*ctx.join(d, algo="fgss"))
        with_sig = ctx.stats()
        b = H.sort_pairs(*ctx.join(d, algo="fgss", tune=H.Tune(flags=H.TUNE_NO_SIGNATURE)))
        without = ctx.stats()
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert with_sig["signature_bits"] == 64
        assert with_sig["signature_tests"] == without["candidates"]
        assert with_sig["candidates"] * 5 < without["candidates"]
        _check(a, O.self_join(c, 0.8), "fgss 50K")
    finally:
        d.free()


@pytest.mark.parametrize("bits", [64, 128, 256, 512, 1024])
@pytest.mark.parametrize("t", [0.6, 0.8, 0.9])
def test_bitmap_small(ctx, small, t, bits):
    c, d = small
    ctx.set_threshold(t)
    want = O.self_join(c, t)
    _check(ctx.join(d, algo="bitmap", tune=H.Tune(bitmap_bits=bits)), want, f"bitmap{bits} {t}")


@pytest.mark.parametrize("scenario", [1, 2])
@pytest.mark.parametrize("threads", [1, 3])
def test_hybrid_small(ctx, small, scenario, threads):
    c, d = small
    ctx.set_threshold(0.7)
    want = O.self_join(c, 0.7)
    tune = H.Tune(verify_scenario=scenario, host_threads=threads, batch_candidates=5000)
    _check(ctx.join(d, algo="hybrid", tune=tune), want, f"hybrid s{scenario} x{threads}")


@pytest.mark.parametrize("threads", [1, 4])
def test_hybrid_cpu_algorithms(ctx, small, threads):
    """The hybrid technique's host filter as hybrid::allpairs and
    hybrid::ppjoin (hybrid_algorithms.hpp:209-387): the oracle's pairs
    either way, and the oracle's counts of postings visited and candidates
    sent to the device (PPJoin's positional filter sends fewer)."""
    c, d = small
    try:
        for t in (0.6, 0.8):
            ctx.set_threshold(t)
            cands = {}
            for name, algo in (("allpairs", O.ALLPAIRS), ("ppjoin", O.PPJOIN)):
                want = O.self_join(c, t, algo=algo)
                ctx.set_host_algorithm(name)
                tune = H.Tune(host_threads=threads, batch_candidates=5000)
                _check(ctx.join(d, algo="hybrid", tune=tune), want, f"hybrid {name} t={t}")
                st = ctx.stats()
                assert st["candidates"] == want.candidates and st["posting_visits"] == want.posting_visits, name
                cands[name] = st["candidates"]
            assert cands["ppjoin"] < cands["allpairs"], t
    finally:
        ctx.set_host_algorithm("allpairs")


def test_block_pairs(ctx, small):
    """join(block*, block*): index over one block, probes from a later block."""
    c, d = small
    ctx.set_threshold(0.6)
    for ib, pb in [((0, 999), (1000, 1999)), ((1000, 1999), (1000, 1999)), ((500, 2500), (2000, 2999))]:
        want = O.self_join_blocks(c, 0.6, ib, pb)
        for algo in ("fgss", "bitmap", "hybrid"):
            _check(ctx.join(d, algo=algo, indexed_range=ib, probe_range=pb), want, f"{algo} {ib} {pb}")


@pytest.mark.parametrize("t", [0.7, 0.8, 0.9])
def test_fgss_config2_full(ctx, t):
    """1,000,000 sets, lognormal sizes (median 10, sigma 0.6, [2,200]),
    Zipf(1.1) over 1M tokens, seed 1: bit-exact against the oracle."""
    c = H.generate_self(H.gen_params(1_000_000))
    d = ctx.upload(c)
    try:
        ctx.set_threshold(t)
        want = O.cpu_multi(c, t, 8)
        _check(ctx.join(d, algo="fgss"), want, f"fgss 1M t={t}")
        st = ctx.stats()
        assert st["tile_plan"] == H.PLAN_TOKEN
        # item-major with its automatic tiles: 4.1 index entries per token at
        # t = 0.7 (balanced tiles), 2.9 and 1.8 above (upload-time tiles)
        _check(ctx.join(d, algo="fgss", tune=H.Tune(flags=H.TUNE_ITEM_MAJOR)), want, f"fgss 1M item-major t={t}")
        sti = ctx.stats()
        assert sti["tile_plan"] == (H.PLAN_BALANCED if t == 0.7 else H.PLAN_UPLOAD)
        for k in ("posting_visits", "candidates", "signature_tests", "index_entries", "probe_prefix_tokens"):
            assert sti[k] == st[k], k
        if t == 0.8:
            _check(ctx.join(d, algo="fgss", tune=H.Tune(flags=H.TUNE_TILE_QUEUE)), want, "fgss 1M tile queue")
            _check(ctx.join(d, algo="fgss", tune=H.Tune(flags=H.TUNE_BALANCED_TILES)), want, "fgss 1M balanced")
        # postings inside the length windows: the same count as the oracle's
        assert st["posting_visits"] == want.posting_visits
        # verified candidates: at most the oracle's (positional filter drops more)
        assert 0 < st["candidates"] <= want.candidates
        sizes, counts = np.unique(c.sizes, return_counts=True)
        mids = sum(int(k) * O.bounds("jaccard", t, int(l))[4] for l, k in zip(sizes, counts))
        assert st["index_entries"] == mids
    finally:
        d.free()


def test_verify_candidates(ctx, small):
    """hs_verify_candidates on a CSR batch (duplicates, ids on both sides of
    the probe) keeps exactly the candidates with |r n s| >= minoverlap."""
    c, d = small
    t = 0.5
    ctx.set_threshold(t)
    rng = np.random.default_rng(0)
    probes = np.arange(100, 3000, 7, dtype=np.int64)
    cands, offs = [], []
    for p in probes:
        cands.extend(rng.integers(max(0, p - 60), min(c.n, p + 20), rng.integers(0, 40)).tolist())
        offs.append(len(cands))
    exp = []
    for p, row in zip(probes, np.split(np.array(cands, dtype=np.uint32), offs[:-1])):
        for s in row:
            ov = O.intersect(c.set(p), c.set(s))
            if ov >= O.bounds("jaccard", t, int(c.sizes[p]), int(c.sizes[s]))[2]:
                exp.append((int(p), int(s), ov))
    exp.sort()
    assert len(exp) > 50
    for scen in (1, 2):
        r, s, ov = H.sort_pairs(*ctx.verify_candidates(d, d, probes, np.array(offs, dtype=np.uint32),
                                                        np.array(cands, dtype=np.uint32), scen))
        assert list(zip(r.tolist(), s.tolist(), ov.tolist())) == exp, f"scenario {scen}"


@pytest.fixture(scope="module")
def rs_small(ctx):
    R, S = H.generate_rs(H.gen_params(1500, size_dist=H.SIZE_UNIFORM, lo=5, hi=60, universe=2000, zipf=0.8, seed=3),
                         H.gen_params(1800, size_dist=H.SIZE_UNIFORM, lo=5, hi=60, universe=2000, zipf=0.8, seed=4),
                         dup_frac=0.3)
    return R, S, ctx.upload(R), ctx.upload(S)


@pytest.mark.parametrize("sim", ["jaccard", "cosine", "dice"])
@pytest.mark.parametrize("algo", ["fgss", "bitmap", "hybrid"])
def test_rs_join_small(ctx, rs_small, sim, algo):
    """R-S (foreign) join: probes of R against an index over S, no id order."""
    R, S, dR, dS = rs_small
    for t in (0.6, 0.8):
        ctx.set_threshold(t, sim)
        want = O.rs_join(R, S, t, sim)
        assert want.count > 50
        _check(ctx.join(dS, dR, algo=algo), want, f"{algo} {sim} {t}")


def test_rs_long_sets_cosine(ctx):
    """The long-set R-S configuration at 1/10 scale: 20,000 x 20,000 sets of
    100-700 tokens over a 50,000-token Zipf(0.8) universe, 20% of S planted
    near-duplicates, cosine 0.8."""
    R, S = H.generate_rs(H.gen_params(20000, size_dist=H.SIZE_UNIFORM, lo=100, hi=700, universe=50000, zipf=0.8,
                                      seed=1),
                         H.gen_params(20000, size_dist=H.SIZE_UNIFORM, lo=100, hi=700, universe=50000, zipf=0.8,
                                      seed=2), dup_frac=0.2, edit_lo=0.05, edit_hi=0.15)
    dR, dS = ctx.upload(R), ctx.upload(S)
    try:
        ctx.set_threshold(0.8, "cosine")
        want = O.rs_join(R, S, 0.8, "cosine")
        assert want.count >= 3000
        _check(ctx.join(dS, dR, algo="fgss"), want, "fgss long cosine")
        # ~58 index entries per token: balanced tiles in 256-lane workgroups
        assert ctx.stats()["tile_plan"] == H.PLAN_BALANCED
        _check(ctx.join(dS, dR, algo="fgss", tune=H.Tune(flags=H.TUNE_TILE_QUEUE)), want, "fgss long cosine queue")
        for bits in (512, 1024, 0):  # 0: the width from the mean set size (1024 bits here)
            _check(ctx.join(dS, dR, algo="bitmap", tune=H.Tune(bitmap_bits=bits)), want, f"bitmap{bits} long cosine")
    finally:
        dR.free()
        dS.free()


@pytest.mark.parametrize("algo", ["fgss", "bitmap"])
def test_join_graph_replays(ctx, small, rs_small, algo):
    """hs_join with HS_TUNE_GRAPH: the first call runs eagerly and captures,
    later calls replay; every replay, with eager joins of other collections
    and thresholds in between (which rewrite the shared staging table,
    signatures and index), equals the oracle.  A new threshold or range
    captures anew."""
    c, d = small
    R, S, dR, dS = rs_small
    g = H.Tune(flags=H.TUNE_GRAPH, bitmap_bits=128 if algo == "bitmap" else 0)
    for t in (0.8, 0.6):
        ctx.set_threshold(t)
        want = O.self_join(c, t)
        for rep in range(4):
            _check(ctx.join(d, algo=algo, tune=g), want, f"{algo} graph t={t} rep {rep}")
            st = ctx.stats()
            assert st["pairs"] == want.count
            # captured on the first call; replayed once the eager joins in
            # between have stopped growing buffers (a growth captures anew)
            if rep == 0:
                assert st["graph"] == 1, rep
            if rep >= 2:
                assert st["graph"] == 2, rep
            # eager work between replays: another collection, another threshold
            ctx.set_threshold(0.7, "cosine")
            ctx.join(dS, dR, algo=algo)
            ctx.set_threshold(t)
        rng = (100, 2500)
        want_b = O.self_join_blocks(c, t, rng, rng)
        for rep in range(2):
            _check(ctx.join(d, algo=algo, tune=g, indexed_range=rng, probe_range=rng), want_b, f"{algo} graph block")
    # R-S through the graph
    ctx.set_threshold(0.8, "cosine")
    want = O.rs_join(R, S, 0.8, "cosine")
    for rep in range(3):
        _check(ctx.join(dS, dR, algo=algo, tune=g), want, f"{algo} graph R-S rep {rep}")
    ctx.set_threshold(0.8)


def test_sets_longer_than_16_bit_fields(ctx):
    """Sets of more than 65,535 tokens: the size and position fields of an
    index record saturate and the kernels read sizes[] instead, the set's
    tokens do not fit the LDS stage, the minoverlap table lives in HBM, the
    balanced plan is not used.  Six near-duplicates of a 70,000-token set
    among small sets, Jaccard 0.8 and 0.9, both techniques, as the oracle."""
    rng = np.random.default_rng(17)
    base = np.sort(rng.choice(np.arange(1, 300_000), size=70_000, replace=False))
    longs = []
    for k in range(6):
        keep = np.sort(rng.choice(base, size=70_000 - 500 * k, replace=False))
        extra = rng.choice(np.arange(300_000, 320_000), size=300 * k, replace=False)
        longs.append(np.unique(np.concatenate([keep, extra])).tolist())
    smalls = [np.sort(rng.choice(np.arange(1, 2000), size=int(rng.integers(3, 12)), replace=False)).tolist()
              for _ in range(300)]
    sets = sorted(smalls + longs, key=lambda s: (len(s), s))
    c = H.SetCollection.from_sets(sets)
    assert int(c.sizes.max()) > 65_535
    d = ctx.upload(c)
    try:
        for t in (0.8, 0.9):
            ctx.set_threshold(t)
            want = O.self_join(c, t)
            assert int((np.asarray(want.overlap) > 65_535).sum()) > 0
            _check(ctx.join(d, algo="fgss"), want, f"long sets fgss t={t}")
            for bits in (256, 1024):
                _check(ctx.join(d, algo="bitmap", tune=H.Tune(bitmap_bits=bits)), want, f"long sets bitmap{bits} t={t}")
    finally:
        d.free()


def _triples_and_longs(n_small, seed):
    """n_small sets of 3 tokens below 70,000 plus three near-duplicate
    50,000-token sets, sorted by (size, tokens)."""
    rng = np.random.default_rng(seed)
    t0 = rng.integers(1, 23_333, size=n_small)
    t1 = t0 + rng.integers(1, 23_333, size=n_small)
    t2 = t1 + rng.integers(1, 23_333, size=n_small)
    tri = np.stack([t0, t1, t2], axis=1)
    tri = tri[np.lexsort((tri[:, 2], tri[:, 1], tri[:, 0]))]
    base = np.sort(rng.choice(np.arange(1, 70_000), size=50_000, replace=False))
    longs = [base]
    for k in (1, 2):
        drop = rng.choice(50_000, size=300 * k, replace=False)
        add = rng.choice(np.setdiff1d(np.arange(1, 70_000), base), size=300 * k, replace=False)
        longs.append(np.sort(np.concatenate([np.delete(base, drop), add])))
    longs.sort(key=lambda a: a.tolist())
    tokens = np.concatenate([tri.reshape(-1)] + longs).astype(np.uint32)
    sizes = np.array([3] * n_small + [50_000] * 3, dtype=np.uint32)
    starts = np.zeros_like(sizes)
    starts[1:] = np.cumsum(sizes[:-1])
    return H.SetCollection(tokens, starts, sizes, int(tokens.max()) + 1)


@pytest.mark.parametrize("n_small,plan", [(3000, "token"), (1_048_600, "item")])
def test_token_plan_word_packing(ctx, n_small, plan):
    """The token-major plan packs an index entry's position, set size, set
    id and key into one 64-bit word.  With 50,000-token sets (16-bit size
    and position fields) among 3,003 sets the fields take 62 bits and the
    plan runs with positions past 10,000; among 1,048,603 sets they would
    take 71 bits and the join takes the item-major plan.  Both as the
    oracle, at Jaccard 0.8 and 0.9."""
    c = _triples_and_longs(n_small, 23)
    d = ctx.upload(c)
    try:
        for t in (0.8, 0.9):
            ctx.set_threshold(t)
            want = O.self_join(c, t)
            assert want.count >= 3
            _check(ctx.join(d, algo="fgss"), want, f"packing n={c.n} t={t}")
            st = ctx.stats()
            assert (st["tile_plan"] == H.PLAN_TOKEN) == (plan == "token"), st["tile_plan"]
    finally:
        d.free()


@pytest.mark.parametrize("algo", ["fgss", "bitmap", "hybrid"])
def test_degenerate_inputs(ctx, algo):
    """Empty collection, one set, 400 identical sets (every pair a result),
    singletons only, threshold 1.0 (exact duplicates), and an empty side of
    an R-S join: every technique, and the co-processing join, as the
    oracle."""
    cases = {
        "empty": [],
        "one": [[3, 9, 12]],
        "identical": [[2, 5, 7, 11, 13]] * 400,
        "singletons": [[i % 50] for i in range(600)],
        "dups_and_near": sorted([[1, 2, 3, 4]] * 30 + [[1, 2, 3, 5]] * 20 + [[1, 2, 3, 4, 6]] * 10,
                                key=lambda s: (len(s), s)),
    }
    for name, sets in cases.items():
        c = H.SetCollection.from_sets(sorted(sets, key=lambda s: (len(s), s)))
        d = ctx.upload(c)
        try:
            for t in (0.5, 0.8, 1.0):
                ctx.set_threshold(t)
                want = O.self_join(c, t)
                _check(ctx.join(d, algo=algo), want, f"{algo} {name} t={t}")
                if algo != "hybrid":
                    (r, s, ov), _ = ctx.join_coprocess(d, algo=algo, cpu_threads=2, cpu_block=7, device_min_claim=64,
                                                       min_device_ms=-1)
                    _check((r, s, ov), want, f"coprocess {algo} {name} t={t}")
        finally:
            d.free()
    # R-S with an empty side
    full = H.SetCollection.from_sets([[1, 2, 3], [1, 2, 4], [2, 3, 4, 5]])
    none = H.SetCollection.from_sets([])
    df, dn = ctx.upload(full), ctx.upload(none)
    try:
        ctx.set_threshold(0.5)
        for idx, prb in ((df, dn), (dn, df)):
            r, s, ov = ctx.join(idx, prb, algo=algo)
            assert r.shape[0] == 0
    finally:
        df.free()
        dn.free()


@pytest.mark.parametrize("algo", ["fgss", "bitmap", "hybrid"])
def test_stage_times(ctx, small, algo):
    """A plain hs_join is timed as a whole (total_ms; the stage times stay 0,
    no marker events between kernels); HS_TUNE_STAGE_TIMES fills index_ms,
    filter_ms and kernel_ms (fgss: from the dispatches' own events) within
    total_ms; the hybrid technique always records its stages.  The pairs
    are the oracle's either way, and a graph replay is timed as a whole."""
    c, d = small
    ctx.set_threshold(0.7)
    want = O.self_join(c, 0.7)
    plain = H.Tune(flags=0)
    staged = H.Tune(flags=H.TUNE_STAGE_TIMES)
    _check(ctx.join(d, algo=algo, tune=plain), want, f"{algo} plain")
    st = ctx.stats()
    assert st["total_ms"] > 0
    if algo == "hybrid":
        assert st["filter_ms"] > 0
    else:
        assert st["index_ms"] == 0 and st["kernel_ms"] == 0 and st["filter_ms"] == 0
    _check(ctx.join(d, algo=algo, tune=staged), want, f"{algo} staged")
    st = ctx.stats()
    assert 0 < st["kernel_ms"] <= st["total_ms"]
    assert 0 < st["filter_ms"] <= st["total_ms"]
    if algo == "fgss":
        assert 0 < st["index_ms"] <= st["total_ms"]
        assert st["kernel_ms"] <= st["filter_ms"]
    # count-only joins after a copying one keep the same timing rules
    assert ctx.join(d, algo=algo, tune=staged, count_only=True) == want.count
    assert ctx.stats()["kernel_ms"] > 0
    if algo != "hybrid":
        g = H.Tune(flags=H.TUNE_GRAPH)
        for _ in range(3):
            assert ctx.join(d, algo=algo, tune=g, count_only=True) == want.count
        st = ctx.stats()
        assert st["graph"] == 2
        assert st["index_ms"] == 0 and st["kernel_ms"] == 0 and 0 < st["filter_ms"] <= st["total_ms"]


def test_token_join_stats_repeat_on_a_tiny_collection(ctx):
    """The work counters of a repeated token-major join over 250 sets (an
    index build of one workgroup) come out the same each time: the build
    clears every counter word, the last statistics slot included, whatever
    its grid.  The sets are equal, so every pair is a candidate and the
    verification's workgroups reach that slot."""
    c = H.SetCollection.from_sets([list(range(1, 9))] * 250)
    d = ctx.upload(c)
    ctx.set_threshold(0.8)
    want = O.self_join(c, 0.8)
    assert want.count == 250 * 249 // 2
    seen = []
    for _ in range(3):
        _check(ctx.join(d, algo="fgss"), want, "fgss equal sets")
        st = ctx.stats()
        seen.append((st["posting_visits"], st["candidates"], st["signature_tests"], st["tile_plan"]))
    assert seen[0][3] == H.PLAN_TOKEN and seen[0][1] >= want.count, seen[0]
    assert seen[1] == seen[0] and seen[2] == seen[0], seen
