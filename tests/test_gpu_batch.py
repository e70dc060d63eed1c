"""Batched joins: one index build and one persistent launch over many
independent self-joins give, join by join, the oracle's result."""
import random

import numpy as np
import pytest

import hyset_amd as H
import oracle as O

pytestmark = pytest.mark.gpu


def make_batch(k, seed):
    rnd = random.Random(seed)
    colls = []
    for j in range(k):
        n = int(round(np.exp(rnd.uniform(np.log(50), np.log(4000)))))
        colls.append(H.generate_self(H.gen_params(n, median=rnd.uniform(4, 20), sigma=0.6, lo=1, hi=120,
                                                  universe=max(n, 20), zipf=rnd.uniform(0.8, 1.3), seed=seed + j)))
    return colls


def split(batch_dc, r, s, ov):
    """Batch ids -> per-join local (r, s, overlap) triples, sorted."""
    jf = batch_dc.join_first
    out = {}
    j_of = np.searchsorted(jf, r, side="right") - 1
    assert np.array_equal(j_of, np.searchsorted(jf, s, side="right") - 1), "pair crosses joins"
    for j in range(len(jf) - 1):
        m = j_of == j
        rr, ss, oo = H.sort_pairs(r[m] - jf[j], s[m] - jf[j], ov[m])
        out[j] = list(zip(rr.tolist(), ss.tolist(), oo.tolist()))
    return out


def expected(colls, t):
    return {j: list(zip(*(x.tolist() for x in (res.r, res.s, res.overlap))))
            for j, res in enumerate(O.self_join(c, t) for c in colls)}


@pytest.fixture(scope="module")
def batch(ctx):
    colls = make_batch(40, 100)
    return colls, ctx.upload_batch(colls)


@pytest.mark.parametrize("t", [0.6, 0.75, 0.9])
def test_batch_equals_per_join_oracle(ctx, batch, t):
    colls, b = batch
    ctx.set_threshold(t)
    want = expected(colls, t)
    assert sum(len(v) for v in want.values()) > 100
    got = split(b, *ctx.join_batch(b))
    assert got == want


def test_batch_tile_sizes_and_grid(ctx, batch):
    colls, b = batch
    ctx.set_threshold(0.75)
    want = expected(colls, 0.75)
    for probes, wpc in ((64, 2), (128, 8), (256, 1)):
        tune = H.Tune(probes_per_tile=probes, tile_tokens=probes * 16, workgroups_per_cu=wpc)
        assert split(b, *ctx.join_batch(b, tune=tune)) == want, (probes, wpc)
    # signature widths and no signature test
    for tune in (H.Tune(signature_bits=128), H.Tune(signature_bits=1024), H.Tune(flags=H.TUNE_NO_SIGNATURE)):
        assert split(b, *ctx.join_batch(b, tune=tune)) == want, (tune.signature_bits, tune.flags)


@pytest.mark.parametrize("t", [0.6, 0.75, 0.9])
def test_batch_threads_per_probe_and_prefix_scheme(ctx, batch, t):
    """The autotuner's knobs give the oracle's pairs: 256-lane workgroups
    (HS_TUNE_NARROW) at 1, 2 and 4 threads per probe, and the index over
    every set's probe prefix (HS_TUNE_INDEX_MAXPREFIX), also replayed from a
    graph; the maxprefix index holds more entries."""
    colls, b = batch
    ctx.set_threshold(t)
    want = expected(colls, t)
    for probes, wpc in ((256, 4), (128, 8), (64, 2)):
        tune = H.Tune(flags=H.TUNE_NARROW, probes_per_tile=probes, tile_tokens=probes * 16, workgroups_per_cu=wpc)
        assert split(b, *ctx.join_batch(b, tune=tune)) == want, ("narrow", probes, wpc)
    mid_entries = (ctx.join_batch(b, count_only=True), ctx.stats()["index_entries"])[1]
    for flags in (H.TUNE_INDEX_MAXPREFIX, H.TUNE_INDEX_MAXPREFIX | H.TUNE_NARROW,
                  H.TUNE_INDEX_MAXPREFIX | H.TUNE_GRAPH):
        for rep in range(2 if flags & H.TUNE_GRAPH else 1):
            assert split(b, *ctx.join_batch(b, tune=H.Tune(flags=flags))) == want, (flags, rep)
    ctx.join_batch(b, tune=H.Tune(flags=H.TUNE_INDEX_MAXPREFIX), count_only=True)
    assert ctx.stats()["index_entries"] > mid_entries


def test_batch_graph_replays(ctx, batch):
    """Replays of the captured graph, with eager joins in between, all equal
    the eager result."""
    colls, b = batch
    ctx.set_threshold(0.75)
    eager = split(b, *ctx.join_batch(b))
    single = ctx.upload(colls[3])
    tune = H.Tune(flags=H.TUNE_GRAPH)
    for rep in range(4):
        assert split(b, *ctx.join_batch(b, tune=tune)) == eager, rep
        ctx.join(single)  # eager work between replays
    ctx.set_threshold(0.6)  # a new threshold re-captures
    assert split(b, *ctx.join_batch(b, tune=tune)) == expected(colls, 0.6)


def test_batch_graph_recaptured_after_buffers_move(ctx, batch):
    """A graph records raw buffer addresses.  After a finer tile plan (a
    larger tile table) and wider signatures reallocate the buffers of the
    batch, the next graph join with the captured configuration captures
    anew instead of replaying, and every join equals the oracle."""
    colls, b = batch
    ctx.set_threshold(0.75)
    want = expected(colls, 0.75)
    g = H.Tune(flags=H.TUNE_GRAPH)
    assert split(b, *ctx.join_batch(b, tune=g)) == want
    finer = H.Tune(probes_per_tile=16, tile_tokens=256, signature_bits=1024)
    assert split(b, *ctx.join_batch(b, tune=finer)) == want
    for rep in range(3):
        assert split(b, *ctx.join_batch(b, tune=g)) == want, rep


def test_autotune_batch(ctx, batch):
    colls, b = batch
    ctx.set_threshold(0.75)
    tune, ms = ctx.autotune_batch(b, reps=1)
    # (all fields 0: the defaults stayed the fastest)
    assert tune.probes_per_tile in (0, 64, 128, 256, 512) and tune.workgroups_per_cu in (0, 1, 2, 4, 8) and ms > 0
    assert tune.probes_per_tile <= (256 if tune.flags & H.TUNE_NARROW else 512)
    assert tune.signature_bits in (0, 64, 128, 256)
    assert split(b, *ctx.join_batch(b, tune=tune)) == expected(colls, 0.75)
    # the tuned configuration as a graph, after the trials moved the buffers
    tune.flags |= H.TUNE_GRAPH
    for rep in range(2):
        assert split(b, *ctx.join_batch(b, tune=tune)) == expected(colls, 0.75), rep
