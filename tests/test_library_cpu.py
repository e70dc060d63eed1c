"""Host-side checks of libhyset_amd.so that need no GPU: the C ABI exports
every declared symbol, the input generator, and the text formats."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import hyset_amd as H

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for hdr in (ROOT / "include").rglob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", hdr.read_text(), flags=re.S)
        names.update(re.findall(r"\b(hs_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(H.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def check_collection(c, lo, hi):
    assert c.sizes.min() >= lo and c.sizes.max() <= hi
    assert np.all(np.diff(c.sizes.astype(np.int64)) >= 0), "sets sorted by size"
    assert c.starts[0] == 0 and np.all(c.starts[1:] == np.cumsum(c.sizes)[:-1])
    for i in range(0, c.n, max(1, c.n // 500)):
        s = c.set(i)
        assert np.all(np.diff(s.astype(np.int64)) > 0), "tokens strictly ascending"
    assert c.tokens.min() >= 1
    assert c.universe == int(c.tokens.max()) + 1


def test_generator_properties_and_determinism():
    p = H.gen_params(20000, median=10, sigma=0.6, lo=2, hi=200, universe=50000, zipf=1.1, seed=1)
    a, b = H.generate_self(p), H.generate_self(p)
    assert np.array_equal(a.tokens, b.tokens) and np.array_equal(a.sizes, b.sizes)
    check_collection(a, 2, 200)
    # ids ranked by ascending frequency
    freq = np.bincount(a.tokens, minlength=a.universe)[1:]
    assert np.all(np.diff(freq) >= 0)
    assert abs(np.median(a.sizes) - 10) <= 1
    p2 = H.gen_params(20000, median=10, sigma=0.6, lo=2, hi=200, universe=50000, zipf=1.1, seed=2)
    assert not np.array_equal(H.generate_self(p2).tokens, a.tokens)


def test_generator_uniform_sizes():
    c = H.generate_self(H.gen_params(5000, size_dist=H.SIZE_UNIFORM, lo=5, hi=50, universe=10000, zipf=1.0))
    check_collection(c, 5, 50)
    assert len(np.unique(c.sizes)) == 46


def test_generator_rs_near_duplicates():
    r, s = H.generate_rs(H.gen_params(500, size_dist=H.SIZE_UNIFORM, lo=100, hi=300, universe=5000, zipf=0.8, seed=1),
                         H.gen_params(500, size_dist=H.SIZE_UNIFORM, lo=100, hi=300, universe=5000, zipf=0.8, seed=2),
                         dup_frac=0.2)
    check_collection(r, 100, 300)
    assert s.sizes.min() >= 80
    rsets = {tuple(r.set(i).tolist()) for i in range(r.n)}
    close = 0
    for j in range(s.n):
        sj = set(s.set(j).tolist())
        best = max(len(sj & set(x)) / len(sj | set(x)) for x in list(rsets)[:500])
        close += best >= 0.7
    assert 80 <= close <= 120  # ~20% planted, 5-15% edits


def test_generator_queries_share_ranking():
    idx, q = H.generate_queries(H.gen_params(5000, universe=20000, seed=1), H.gen_params(300, universe=20000, seed=2))
    check_collection(idx, 2, 200)
    # query tokens unseen in the index get ids above the index universe
    assert q.tokens.max() >= idx.universe - 1
    assert np.isin(q.tokens[q.tokens < idx.universe], np.unique(idx.tokens)).all()


def test_text_round_trip(tmp_path):
    c = H.generate_self(H.gen_params(300, size_dist=H.SIZE_UNIFORM, lo=1, hi=9, universe=100, zipf=1.0))
    path = tmp_path / "c.txt"
    H.write_collection_text(path, c)
    d = H.read_collection_text(path)
    assert np.array_equal(c.tokens, d.tokens) and np.array_equal(c.sizes, d.sizes)
    # empty lines are skipped, as the reference's loader does
    path.write_text("1 2 3\n\n4 5\n\n")
    e = H.read_collection_text(path)
    assert e.n == 2 and e.sizes.tolist() == [3, 2]
    pairs = tmp_path / "p.txt"
    H.write_pairs_text(pairs, [3, 4], [1, 2], None)
    assert pairs.read_text() == "3 1\n4 2\n"
    H.write_pairs_text(pairs, [3], [1], [7])
    assert pairs.read_text() == "3 1 7\n"


def test_pairs_text_writer_formatting(tmp_path):
    """The buffered writer's own decimal formatting against Python's, over
    every digit count, the extremes of uint32 and more than one 1 MiB buffer."""
    rng = np.random.default_rng(3)
    edge = np.array([0, 1, 9, 10, 99, 100, 4294967295, 4294967294, 1000000000, 999999999], dtype=np.uint32)
    r = np.concatenate([edge, rng.integers(0, 2**32, 60000, dtype=np.uint64).astype(np.uint32)])
    s = np.concatenate([edge[::-1], (rng.integers(0, 2**32, 60000, dtype=np.uint64) >> rng.integers(0, 32, 60000,
                                                                                              dtype=np.uint64)
                                     ).astype(np.uint32)])
    ov = (r % 1000).astype(np.uint32)
    p = tmp_path / "p.txt"
    H.write_pairs_text(p, r, s, ov)
    assert p.read_text() == "".join(f"{a} {b} {c}\n" for a, b, c in zip(r.tolist(), s.tolist(), ov.tolist()))
    H.write_pairs_text(p, r, s)
    assert p.read_text() == "".join(f"{a} {b}\n" for a, b in zip(r.tolist(), s.tolist()))
    H.write_pairs_text(p, [], [])
    assert p.read_text() == ""
    with pytest.raises(H.HysetError):
        H.write_pairs_text(tmp_path / "no" / "such" / "dir.txt", [1], [2])


@pytest.mark.parametrize("sim", ["jaccard", "cosine", "dice"])
def test_host_join_ppjoin_matches_oracle(sim):
    """hs_join_host with each CPU filtering algorithm (host::allpairs /
    host::ppjoin): the oracle's pairs and overlaps, and the oracle's count of
    postings visited and of candidates -- PPJoin's positional filter keeps
    fewer candidates (the oracle restates host_algorithms.hpp:131-250)."""
    import oracle as O
    c = H.generate_self(H.gen_params(3000, size_dist=H.SIZE_UNIFORM, lo=2, hi=40, universe=800, zipf=1.0, seed=5))
    with H.Context(-1) as ctx:
        d = ctx.upload(c)
        for t in (0.6, 0.8, 0.9):
            ctx.set_threshold(t, sim)
            fewer = False
            for name, algo in (("allpairs", O.ALLPAIRS), ("ppjoin", O.PPJOIN)):
                want = O.self_join(c, t, sim, algo=algo)
                ctx.set_host_algorithm(name)
                for threads in (1, 3):
                    r, s, ov = H.sort_pairs(*ctx.join_host(d, threads=threads))
                    assert np.array_equal(r, want.r) and np.array_equal(s, want.s), (name, t)
                    assert np.array_equal(ov, want.overlap)
                    st = ctx.stats()
                    assert st["posting_visits"] == want.posting_visits, (name, t)
                    assert st["candidates"] == want.candidates, (name, t)
                if name == "allpairs":
                    all_cands = want.candidates
                else:
                    fewer = want.candidates < all_cands
            assert fewer, t
        ctx.set_host_algorithm("allpairs")
        with pytest.raises(H.HysetError):
            ctx.set_host_algorithm(7)


def test_host_join_matches_oracle():
    """hs_join_host on a host-only context: the library's own CPU partition."""
    import oracle as O
    c = H.generate_self(H.gen_params(3000, size_dist=H.SIZE_UNIFORM, lo=2, hi=40, universe=800, zipf=1.0, seed=5))
    with H.Context(-1) as ctx:
        d = ctx.upload(c)
        for sim in ("jaccard", "cosine", "dice"):
            for t in (0.6, 0.8, 0.9):
                ctx.set_threshold(t, sim)
                want = O.self_join(c, t, sim)
                for threads in (1, 4):
                    r, s, ov = H.sort_pairs(*ctx.join_host(d, threads=threads))
                    assert np.array_equal(r, want.r) and np.array_equal(s, want.s)
                    assert np.array_equal(ov, want.overlap)
        r2, s2 = H.generate_rs(H.gen_params(700, size_dist=H.SIZE_UNIFORM, lo=5, hi=40, universe=500, zipf=0.8, seed=3),
                               H.gen_params(800, size_dist=H.SIZE_UNIFORM, lo=5, hi=40, universe=500, zipf=0.8, seed=4),
                               dup_frac=0.3)
        dr, ds = ctx.upload(r2), ctx.upload(s2)
        ctx.set_threshold(0.7, "cosine")
        want = O.rs_join(r2, s2, 0.7, "cosine")
        r, s, ov = H.sort_pairs(*ctx.join_host(ds, dr, threads=3))
        assert np.array_equal(r, want.r) and np.array_equal(s, want.s) and np.array_equal(ov, want.overlap)
        # device entry points refuse a host-only context
        try:
            ctx.join(d)
            raise AssertionError("hs_join ran on a host-only context")
        except H.HysetError as e:
            assert e.code == H.HS_E_INVALID
