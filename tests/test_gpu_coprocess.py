"""One join shared by the device and host threads (hs_join_coprocess): the
work-sharing queue and the static split give the oracle's pairs for every
way the probe ids are divided, self-join and R-S, and the per-side counts
add up."""
import numpy as np
import pytest

import hyset_amd as H
import oracle as O

pytestmark = pytest.mark.gpu


def _triples(r, s, ov):
    return sorted(zip(r.tolist(), s.tolist(), ov.tolist()))


def _want(res):
    return sorted(zip(res.r.tolist(), res.s.tolist(), res.overlap.tolist()))


@pytest.fixture(scope="module")
def col(ctx):
    c = H.generate_self(H.gen_params(20000, median=10, sigma=0.6, lo=2, hi=200, universe=20000, zipf=1.1, seed=5))
    d = ctx.upload(c)
    yield c, d
    d.free()


def _check_stats(info, n, count):
    assert info["device_probes"] + info["host_probes"] == n
    assert info["device_pairs"] + info["host_pairs"] == count
    assert info["split"] == n - info["device_probes"]


@pytest.mark.parametrize("algo", ["fgss", "bitmap"])
@pytest.mark.parametrize("kw", [dict(cpu_threads=4, min_device_ms=-1),
                                dict(cpu_threads=1, cpu_block=1, device_min_claim=256, min_device_ms=-1),
                                dict(cpu_threads=3, cpu_block=64, device_min_claim=1000, min_device_ms=-1)])
def test_queue_self_join(ctx, col, algo, kw):
    c, d = col
    for t in (0.6, 0.8):
        ctx.set_threshold(t)
        want = O.self_join(c, t)
        (r, s, ov), info = ctx.join_coprocess(d, algo=algo, **kw)
        assert _triples(r, s, ov) == _want(want), (t, kw)
        _check_stats(info, c.n, want.count)
        if kw.get("device_min_claim") == 256:
            assert info["device_claims"] >= 2


@pytest.mark.parametrize("share", [0.0, 0.05, 0.5, 1.0])
def test_static_split_self_join(ctx, col, share):
    c, d = col
    ctx.set_threshold(0.7)
    want = O.self_join(c, 0.7)
    (r, s, ov), info = ctx.join_coprocess(d, mode="static", cpu_share=share, cpu_threads=4)
    assert _triples(r, s, ov) == _want(want), share
    _check_stats(info, c.n, want.count)
    assert info["split"] == round(share * c.n)
    assert info["device_claims"] == (0 if share == 1.0 else 1)


@pytest.mark.parametrize("mode,kw", [("static", dict(cpu_share=0.5, cpu_threads=4)),
                                     ("queue", dict(cpu_threads=3, cpu_block=64, device_min_claim=1000,
                                                    min_device_ms=-1))])
def test_host_side_ppjoin(ctx, col, mode, kw):
    """The host threads filter as PPJoin (positions from the device-built
    index's entries): the same pairs as the oracle, the host side's share
    still present."""
    c, d = col
    try:
        ctx.set_host_algorithm("ppjoin")
        for t in (0.6, 0.8):
            ctx.set_threshold(t)
            want = O.self_join(c, t, algo=O.PPJOIN)
            (r, s, ov), info = ctx.join_coprocess(d, mode=mode, **kw)
            assert _triples(r, s, ov) == _want(want), (mode, t)
            _check_stats(info, c.n, want.count)
            if mode == "static":
                assert info["host_probes"] == round(0.5 * c.n)
    finally:
        ctx.set_host_algorithm("allpairs")


def test_break_even_gate(ctx, col):
    """Queue mode with the gate: a join with no measured device-alone time
    runs on the device alone and records it; one below min_device_ms stays
    there; above it the next call shares once, and the one after goes by the
    two measured wall times.  Every answer is the oracle's."""
    c, d = col
    ctx.set_threshold(0.7)
    want = _want(O.self_join(c, 0.7))
    for _ in range(2):  # the default gate (50 ms): this 20K-set join never shares
        (r, s, ov), info = ctx.join_coprocess(d, cpu_threads=2)
        assert _triples(r, s, ov) == want
        assert info["gated"] == 1 and info["host_probes"] == 0 and info["device_probes"] == c.n
        assert info["device_alone_ms"] > 0
    kw = dict(cpu_threads=3, cpu_block=64, min_device_ms=1e-6)
    (r, s, ov), i1 = ctx.join_coprocess(d, **kw)  # a new join key: measured alone first
    assert _triples(r, s, ov) == want and i1["gated"] == 1
    (r, s, ov), i2 = ctx.join_coprocess(d, **kw)  # long enough by this gate: shares once
    assert _triples(r, s, ov) == want and i2["gated"] == 0 and i2["shared_ms"] > 0
    _check_stats(i2, c.n, len(want))
    (r, s, ov), i3 = ctx.join_coprocess(d, **kw)  # by the measured times
    assert _triples(r, s, ov) == want
    assert i3["gated"] == int(i2["shared_ms"] >= i2["device_alone_ms"]), (i2, i3)


def test_rs_join_both_modes(ctx):
    R, S = H.generate_rs(H.gen_params(3000, size_dist=H.SIZE_UNIFORM, lo=5, hi=60, universe=2000, zipf=0.8, seed=3),
                         H.gen_params(2500, size_dist=H.SIZE_UNIFORM, lo=5, hi=60, universe=2000, zipf=0.8, seed=4),
                         dup_frac=0.3)
    dR, dS = ctx.upload(R), ctx.upload(S)
    try:
        for sim in ("jaccard", "cosine"):
            ctx.set_threshold(0.7, sim)
            want = O.rs_join(R, S, 0.7, sim)
            assert want.count > 100
            for kw in (dict(mode="queue", cpu_threads=2, cpu_block=8, device_min_claim=300, min_device_ms=-1),
                       dict(mode="static", cpu_share=0.3, cpu_threads=2)):
                (r, s, ov), info = ctx.join_coprocess(dS, dR, **kw)
                assert _triples(r, s, ov) == _want(want), (sim, kw)
                _check_stats(info, R.n, want.count)
    finally:
        dR.free()
        dS.free()


def test_count_only_then_copy_and_dense_host_share(ctx):
    """Many pairs on both sides: the host's pairs land behind the device's in
    the context's result, where hs_copy_last_result and the stream sink find
    them."""
    c = H.generate_self(H.gen_params(3000, size_dist=H.SIZE_UNIFORM, lo=4, hi=6, universe=12, zipf=0.5, seed=9))
    d = ctx.upload(c)
    try:
        ctx.set_threshold(0.5)
        want = O.self_join(c, 0.5)
        n, info = ctx.join_coprocess(d, mode="static", cpu_share=0.6, cpu_threads=4, count_only=True)
        assert n == want.count and info["host_pairs"] > 100000 and info["device_pairs"] > 100000
        r, s, ov = ctx.last_result(n)
        assert _triples(r, s, ov) == _want(want)
        # device pairs first, then the host's (ids below the split)
        g = info["device_pairs"]
        assert np.all(r[:g] >= info["split"]) and np.all(r[g:] < info["split"])
        got = []
        ctx.stream_last_result(lambda a, b, o: got.append((a, b, o)), chunk=100000)
        assert np.array_equal(np.concatenate([x[0] for x in got]), r)
        assert np.array_equal(np.concatenate([x[2] for x in got]), ov)
        # an ordinary join afterwards is unaffected
        assert ctx.join(d, algo="fgss", count_only=True) == want.count
    finally:
        d.free()


@pytest.mark.parametrize("policy", ["cpu,gpu", "gpu,cpu"])
@pytest.mark.parametrize("algo", ["fgss", "bitmap", "hybrid"])
def test_static_policies(ctx, col, policy, algo):
    """The --policy pairs of dichotomy.cu:179-258 with every device technique
    (hybrid: the cpu-gpu side, its stripes filtered on the host and verified
    on the device): Rs = the probe sets below share * n goes to the first
    side, Rl to the second; the pairs are the oracle's for every share."""
    c, d = col
    tune = H.Tune(host_threads=2) if algo == "hybrid" else None
    for t, share in ((0.6, 0.3), (0.8, 0.0), (0.8, 1.0), (0.7, 0.75)):
        ctx.set_threshold(t)
        want = O.self_join(c, t)
        (r, s, ov), info = ctx.join_coprocess(d, mode="static", policy=policy, algo=algo, cpu_share=share,
                                              cpu_threads=3, tune=tune)
        assert _triples(r, s, ov) == _want(want), (t, share)
        cut = round(share * c.n)
        assert info["split"] == cut
        assert info["device_probes"] + info["host_probes"] == c.n
        assert info["device_pairs"] + info["host_pairs"] == want.count
        dev_part = cut if policy == "gpu,cpu" else c.n - cut
        assert info["device_probes"] == dev_part
        assert info["device_claims"] == (1 if dev_part else 0)
        if policy == "gpu,cpu":
            # the device's pairs have their larger id in Rs, the host's in Rl
            g = info["device_pairs"]
            assert np.all(r[:g] < cut) and np.all(r[g:] >= cut)


@pytest.mark.parametrize("host_algo", ["allpairs", "ppjoin"])
@pytest.mark.parametrize("filter_threads", [1, 3])
def test_queue_cooperative_hybrid(ctx, col, host_algo, filter_threads):
    """queue.cu's --cooperative device consumer: the device side's claims
    are filtered on the host (AllPairs or PPJoin, 1 or 3 filter threads) and
    verified on the device while the host threads join from the bottom."""
    c, d = col
    try:
        ctx.set_host_algorithm(host_algo)
        for t in (0.6, 0.8):
            ctx.set_threshold(t)
            want = O.self_join(c, t)
            (r, s, ov), info = ctx.join_coprocess(d, algo="hybrid", cpu_threads=2, cpu_block=32,
                                                  device_min_claim=500, min_device_ms=-1,
                                                  tune=H.Tune(host_threads=filter_threads))
            assert _triples(r, s, ov) == _want(want), (t, host_algo)
            _check_stats(info, c.n, want.count)
            assert info["device_claims"] >= 1
    finally:
        ctx.set_host_algorithm("allpairs")


def test_rs_join_policies_hybrid(ctx):
    R, S = H.generate_rs(H.gen_params(3000, size_dist=H.SIZE_UNIFORM, lo=5, hi=60, universe=2000, zipf=0.8, seed=3),
                         H.gen_params(2500, size_dist=H.SIZE_UNIFORM, lo=5, hi=60, universe=2000, zipf=0.8, seed=4),
                         dup_frac=0.3)
    dR, dS = ctx.upload(R), ctx.upload(S)
    try:
        ctx.set_threshold(0.7, "cosine")
        want = _want(O.rs_join(R, S, 0.7, "cosine"))
        for kw in (dict(mode="static", policy="gpu,cpu", algo="hybrid", cpu_share=0.4),
                   dict(mode="static", policy="gpu,cpu", algo="fgss", cpu_share=0.4),
                   dict(mode="queue", algo="hybrid", cpu_block=16, device_min_claim=300, min_device_ms=-1)):
            (r, s, ov), info = ctx.join_coprocess(dS, dR, cpu_threads=2, **kw)
            assert _triples(r, s, ov) == want, kw
            assert info["device_probes"] + info["host_probes"] == R.n
    finally:
        dR.free()
        dS.free()


def test_bad_options(ctx, col):
    _, d = col
    ctx.set_threshold(0.8)
    with pytest.raises(H.HysetError):
        ctx.join_coprocess(d, algo=7)
    with pytest.raises(H.HysetError):
        ctx.join_coprocess(d, mode="static", cpu_share=1.5)
    with pytest.raises(H.HysetError):
        ctx.join_coprocess(d, mode=7)
    with pytest.raises(H.HysetError):
        ctx.join_coprocess(d, mode="static", policy=5)
    with pytest.raises(H.HysetError):  # the queue's device claims from the large sets
        ctx.join_coprocess(d, mode="queue", policy="gpu,cpu")
