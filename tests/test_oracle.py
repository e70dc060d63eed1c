"""The oracle pinned: against the golden fixtures, an independent Python
brute force, its own brute force, and the special cases of the domain."""
from pathlib import Path

import numpy as np
import pytest

import hyset_amd as H
import oracle as O
from tests.golden import make_fixtures as F

GOLD = Path(__file__).resolve().parent / "golden"


def rows(res):
    return list(zip(res.r.tolist(), res.s.tolist(), res.overlap.tolist()))


def test_config1_generator_digest():
    assert F.collection_digest(F.config1()) == (GOLD / "config1.sha256").read_text().strip()


def test_neardup_generator_digest():
    assert F.collection_digest(F.neardup()) == (GOLD / "neardup.sha256").read_text().strip()


@pytest.mark.parametrize("t", F.CONFIG1_T)
def test_config1_fixture(t):
    c = F.config1()
    want = F.read_pairs(GOLD / f"config1_t{t}.pairs")
    assert rows(O.self_join(c, t)) == want
    assert rows(O.self_join(c, t, algo=O.PPJOIN)) == want
    assert rows(O.cpu_multi(c, t, 4)) == want


@pytest.mark.parametrize("t", F.NEARDUP_T)
def test_neardup_fixture(t):
    c = F.neardup()
    want = F.read_pairs(GOLD / f"neardup_t{t}.pairs")
    assert len(want) > 500
    assert rows(O.self_join(c, t)) == want
    assert rows(O.self_join(c, t, algo=O.PPJOIN)) == want
    assert rows(O.cpu_multi(c, t, 3)) == want


@pytest.mark.parametrize("sim,t", F.SPECIAL_CASES)
def test_special_fixture(sim, t):
    c = H.read_collection_text(GOLD / "special.txt")
    want = F.read_pairs(GOLD / f"special_t{t}_{sim}.pairs")
    assert rows(O.self_join(c, t, sim)) == want
    assert rows(O.bruteforce_self(c, t, sim)) == want


def test_special_fixture_rounding_cases():
    """J = 0.8 exactly at sums 63 and 126 is dropped by the reference bounds
    and kept by the exact ones; J = 0.9 exactly at sum 38 likewise."""
    c = H.read_collection_text(GOLD / "special.txt")
    for t, sizes in ((0.8, (31, 32)), (0.8, (63, 63)), (0.9, (19, 19))):
        ref = O.self_join(c, t)
        ex = O.self_join(c, t, mode=O.BOUNDS_EXACT)
        got_ref = {(int(c.sizes[s]), int(c.sizes[r])) for r, s in zip(ref.r, ref.s)}
        got_ex = {(int(c.sizes[s]), int(c.sizes[r])) for r, s in zip(ex.r, ex.s)}
        assert sizes not in got_ref and sizes in got_ex


def small_random(n, lo, hi, U, zipf, seed):
    return H.generate_self(H.gen_params(n, size_dist=H.SIZE_UNIFORM, lo=lo, hi=hi, universe=U, zipf=zipf,
                                        seed=seed))


@pytest.mark.parametrize("sim", ["jaccard", "cosine", "dice"])
@pytest.mark.parametrize("mode", [O.BOUNDS_REFERENCE, O.BOUNDS_EXACT])
def test_allpairs_equals_bruteforce(sim, mode):
    c = small_random(1500, 1, 30, 300, 1.0, 7)
    for t in (0.5, 0.65, 0.8, 0.9, 1.0):
        bf = rows(O.bruteforce_self(c, t, sim, mode))
        assert rows(O.self_join(c, t, sim, mode)) == bf
        assert rows(O.self_join(c, t, sim, mode, algo=O.PPJOIN)) == bf


def test_block_join_equals_bruteforce_restriction():
    c = small_random(1200, 2, 20, 200, 1.0, 9)
    bf = O.bruteforce_self(c, 0.6)
    for ib, pb in (((0, 399), (400, 1199)), ((300, 700), (600, 1000)), ((0, 1199), (0, 1199))):
        got = rows(O.self_join_blocks(c, 0.6, ib, pb))
        want = [(r, s, o) for r, s, o in rows(bf) if ib[0] <= s <= ib[1] and pb[0] <= r <= pb[1]]
        assert got == want


@pytest.mark.parametrize("sim", ["jaccard", "cosine", "dice"])
def test_rs_join_equals_bruteforce(sim):
    r, s = H.generate_rs(H.gen_params(600, size_dist=H.SIZE_UNIFORM, lo=5, hi=40, universe=500, zipf=0.8, seed=3),
                         H.gen_params(700, size_dist=H.SIZE_UNIFORM, lo=5, hi=40, universe=500, zipf=0.8, seed=4),
                         dup_frac=0.3)
    for t in (0.6, 0.8):
        bf = O.bruteforce_rs(r, s, t, sim)
        assert bf.count > 20
        assert rows(O.rs_join(r, s, t, sim)) == rows(bf)


def test_special_cases():
    # identical sets always pair; disjoint never; quadratic duplicate groups
    dup = H.SetCollection.from_sets([[1, 2, 3]] * 50 + [[4, 5, 6, 7]] * 30)
    res = O.self_join(dup, 0.95)
    assert res.count == 50 * 49 // 2 + 30 * 29 // 2
    assert rows(res) == rows(O.bruteforce_self(dup, 0.95))
    disjoint = H.SetCollection.from_sets([[i * 3, i * 3 + 1, i * 3 + 2] for i in range(100)])
    assert O.self_join(disjoint, 0.1).count == 0
    empty = H.SetCollection.from_sets([])
    assert O.self_join(empty, 0.8).count == 0
    one = H.SetCollection.from_sets([[5]])
    assert O.self_join(one, 0.8).count == 0


def test_intersect_and_merge_path_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(300):
        a = np.unique(rng.integers(0, 60, rng.integers(0, 25)))
        b = np.unique(rng.integers(0, 60, rng.integers(0, 25)))
        assert O.intersect(a, b) == len(set(a.tolist()) & set(b.tolist()))
        # merge with ties taking b first: element tagged (value, 0) for b, (value, 1) for a
        merged = sorted([(int(x), 1) for x in a] + [(int(x), 0) for x in b])
        for d in range(len(merged) + 1):
            assert O.merge_path(a, b, d) == sum(1 for v, tag in merged[:d] if tag == 1)


def test_oracle_counters():
    c = small_random(800, 2, 20, 200, 1.0, 2)
    res = O.self_join(c, 0.6)
    assert res.candidates >= res.count
    assert res.posting_visits >= res.candidates
    mids = sum(O.bounds("jaccard", 0.6, int(l))[4] for l in c.sizes)
    assert res.index_entries == mids


def test_oracle_intervals():
    """cpu_single reports the reference's two intervals (index build, then
    the join proper) on a steady clock; cpu_multi the join's wall time and
    its slowest worker's index build."""
    c = small_random(3000, 2, 40, 500, 1.0, 3)
    single = O.self_join(c, 0.6)
    multi = O.cpu_multi(c, 0.6, 3)
    assert single.count == multi.count
    assert single.index_ms > 0 and single.join_ms > 0
    assert multi.join_ms > 0 and 0 < multi.index_ms <= multi.join_ms
