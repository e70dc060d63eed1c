"""The result sink: pairs leave the device in pinned chunks, into the
caller's arrays (hs_copy_last_result), to a streaming callback
(hs_stream_last_result) or into a text file (hs_write_last_result_text).
Chunked transfers are checked against the oracle and against a transfer in
one chunk of the same device buffer, over results of one and of several
chunks."""
import numpy as np
import pytest

import hyset_amd as H
import oracle as O

pytestmark = pytest.mark.gpu


def _dense(n, seed):
    # sets of 4-6 tokens over 12: a few percent of all pairs reach J >= 0.5
    return H.generate_self(H.gen_params(n, size_dist=H.SIZE_UNIFORM, lo=4, hi=6, universe=12, zipf=0.5, seed=seed))


def _stream(ctx, chunk, overlap=True):
    got, sizes = [], []

    def take(r, s, ov):
        got.append((r, s, ov))
        sizes.append(r.shape[0])

    ctx.stream_last_result(take, chunk=chunk, overlap=overlap)
    if not got:
        return np.empty(0, np.uint32), np.empty(0, np.uint32), None, sizes
    r = np.concatenate([g[0] for g in got])
    s = np.concatenate([g[1] for g in got])
    ov = np.concatenate([g[2] for g in got]) if overlap else None
    return r, s, ov, sizes


def test_stream_chunks_match_copy_and_oracle(ctx):
    c = H.generate_self(H.gen_params(3000, size_dist=H.SIZE_UNIFORM, lo=3, hi=12, universe=300, zipf=1.0, seed=8))
    d = ctx.upload(c)
    try:
        ctx.set_threshold(0.5)
        want = O.self_join(c, 0.5)
        assert want.count > 1000
        r, s, ov = ctx.join(d, algo="fgss")
        assert sorted(zip(r.tolist(), s.tolist(), ov.tolist())) == \
            sorted(zip(want.r.tolist(), want.s.tolist(), want.overlap.tolist()))
        for chunk in (1, 7, 1000, 0, 1 << 30):
            gr, gs, gov, sizes = _stream(ctx, chunk)
            assert np.array_equal(gr, r) and np.array_equal(gs, s) and np.array_equal(gov, ov), chunk
            step = chunk if 0 < chunk < r.shape[0] else r.shape[0]
            assert sizes[:-1] == [step] * (len(sizes) - 1) and 0 < sizes[-1] <= step, chunk
        gr, gs, gov, _ = _stream(ctx, 333, overlap=False)
        assert gov is None and np.array_equal(gr, r) and np.array_equal(gs, s)
    finally:
        d.free()


def test_multi_chunk_copy_and_text(ctx, tmp_path):
    """A result of several default chunks (262144 pairs each): the double
    buffered copy equals the same buffer moved in one chunk, and the
    streamed text file equals the formatted copy."""
    c = _dense(3000, 9)
    d = ctx.upload(c)
    try:
        ctx.set_threshold(0.5)
        r, s, ov = ctx.join(d, algo="fgss")
        assert r.shape[0] > 2 * 262144 + 1000
        gr, gs, gov, sizes = _stream(ctx, 1 << 30)
        assert len(sizes) == 1
        assert np.array_equal(gr, r) and np.array_equal(gs, s) and np.array_equal(gov, ov)
        # the pairs themselves: the bitmap technique finds the same set
        rb, sb, ob = ctx.join(d, algo="bitmap")
        assert np.array_equal(np.sort(r.astype(np.uint64) << 32 | s), np.sort(rb.astype(np.uint64) << 32 | sb))
        ctx.join(d, algo="fgss", count_only=True)
        p = tmp_path / "pairs.txt"
        ctx.write_last_result_text(p)
        r2, s2, _ = ctx.last_result(r.shape[0])
        lines = p.read_text().splitlines()
        assert len(lines) == r2.shape[0]
        assert lines[:3] + lines[-3:] == [f"{a} {b}" for a, b in zip(np.r_[r2[:3], r2[-3:]].tolist(),
                                                                         np.r_[s2[:3], s2[-3:]].tolist())]
        got = np.array(p.read_text().split(), dtype=np.uint64).reshape(-1, 2)
        assert np.array_equal(got[:, 0], r2) and np.array_equal(got[:, 1], s2)
        ctx.write_last_result_text(p, overlap=True)
        got = np.array(p.read_text().split(), dtype=np.uint64).reshape(-1, 3)
        assert got.shape[0] == r2.shape[0] and np.array_equal(got[:, 2], ctx.last_result(r.shape[0])[2])
    finally:
        d.free()


def test_sink_stop_and_empty(ctx, tmp_path):
    c = _dense(3000, 10)
    d = ctx.upload(c)
    try:
        ctx.set_threshold(0.5)
        n = ctx.join(d, algo="fgss", count_only=True)
        calls = []
        with pytest.raises(H.HysetError):
            ctx.stream_last_result(lambda r, s, ov: calls.append(1) or len(calls) == 2, chunk=1000)
        assert len(calls) == 2
        # the context is still usable, and the result is still there
        r, s, ov = ctx.last_result(n)
        gr, gs, gov, _ = _stream(ctx, 4096)
        assert np.array_equal(gr, r) and np.array_equal(gs, s) and np.array_equal(gov, ov)
        # an empty result reaches no sink call and writes an empty file
        ctx.set_threshold(1.0)
        e = H.generate_self(H.gen_params(500, size_dist=H.SIZE_UNIFORM, lo=5, hi=5, universe=100000, seed=3))
        de = ctx.upload(e)
        try:
            assert ctx.join(de, algo="fgss", count_only=True) == 0
            gr, _, _, sizes = _stream(ctx, 0)
            assert sizes == [] and gr.shape[0] == 0
            p = tmp_path / "empty.txt"
            ctx.write_last_result_text(p)
            assert p.read_text() == ""
        finally:
            de.free()
    finally:
        d.free()
