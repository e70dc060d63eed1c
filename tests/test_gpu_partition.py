"""Partition pairs of one self-join (cost::run, include/hyset/cost.hpp:28-58,
and its execution, src/cost_model.cu:150-262): the device's two-range
features against the oracle's restatement of statistics::extract(left,
right), the technique choice against the oracle's cost::run on the same
features, and any plan of partition pairs giving exactly the oracle's join."""
import numpy as np
import pytest

import hyset_amd as H
import oracle as O
from oracle import cost as OC

pytestmark = pytest.mark.gpu


def _rows(r, s, ov):
    return sorted(zip(r.tolist(), s.tolist(), ov.tolist()))


def _want(res):
    return sorted(zip(res.r.tolist(), res.s.tolist(), res.overlap.tolist()))


@pytest.fixture(scope="module")
def col(ctx):
    c = H.generate_self(H.gen_params(20000, median=10, sigma=0.6, lo=2, hi=200, universe=20000, zipf=1.1, seed=5))
    d = ctx.upload(c)
    yield c, d
    d.free()


@pytest.mark.parametrize("sim,t", [("jaccard", 0.8), ("jaccard", 0.6), ("cosine", 0.75)])
def test_features_pair_matches_oracle(ctx, col, sim, t):
    c, d = col
    ctx.set_threshold(t, sim)
    for left, right in (((0, 4999), (5000, 19999)), ((0, 0), (1, 1)), ((100, 2000), (2001, 2001)),
                        ((7000, 9999), (12000, 19999)), ((0, 19998), (19999, 19999))):
        got = ctx.features_pair(d, left, right)
        want = OC.features_pair(c, t, left, right, sim=sim, universe=c.universe)
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-6, err_msg=str((left, right)))


def test_cost_run_matches_oracle(ctx, col):
    """hs_cost_run's features are hs_features / hs_features_pair (checked
    against the oracle), its costs the oracle's linear costs of those
    features, its choice the oracle's cost::run; a disabled technique is
    never chosen, and the wrapped square changes the choice where it
    should."""
    c, d = col
    ctx.set_threshold(0.8)
    parts = H.partition_even(c.n, 1500, 4)
    assert len(parts) == 4
    rng = np.random.default_rng(2)
    for trial in range(4):
        coefs = rng.normal(0, 1, (4, 9)) * np.array([50, 1, 1, 1, 1e-3, 1e-9, 1e-6, 1e-4, 0.1])
        enabled = (1, 1, 1, 1) if trial < 2 else (1, 0, 1, 1)
        ij, tech, feats, costs = ctx.cost_run(d, parts, coefs, enabled=enabled, bitmap_bits=128)
        assert len(ij) == 10 and tech.shape == (10,)
        for q, (i, j) in enumerate(ij):
            if i == j:
                want_f = OC.features(c, 0.8, parts[i][0], parts[i][1], universe=c.universe)
            else:
                want_f = OC.features_pair(c, 0.8, parts[i], parts[j], universe=c.universe)
            np.testing.assert_allclose(feats[q], want_f, rtol=2e-6, atol=1e-6)
            for k in range(4):
                if enabled[k]:
                    assert costs[q, k] == pytest.approx(OC.linear_cost(coefs[k], feats[q], 128), rel=1e-12)
                else:
                    assert costs[q, k] == -1
        assert list(tech) == OC.cost_run(feats, coefs, enabled, 128)
        if not enabled[H.TECH_BITMAP]:
            assert H.TECH_BITMAP not in tech
    # size^2 alone decides: exact squares rank by size, wrapped ones as the reference's uint product
    big = H.partition_even(c.n, 10000, 2)  # two partitions of 10,000 sets: pair sizes 1e4, 2e4
    coefs = np.zeros((4, 9))
    coefs[H.TECH_FGSS, 5] = 1.0
    coefs[H.TECH_CPU, 0] = 2.0e8
    ij, tech, feats, _ = ctx.cost_run(d, big, coefs, enabled=(1, 0, 0, 1))
    assert list(tech) == OC.cost_run(feats, coefs, (1, 0, 0, 1), 64)
    ij, tech_w, feats, _ = ctx.cost_run(d, big, coefs, enabled=(1, 0, 0, 1), wrap_square=True)
    assert list(tech_w) == OC.cost_run(feats, coefs, (1, 0, 0, 1), 64, wrap_square=True)


@pytest.mark.parametrize("t", [0.6, 0.8])
def test_any_plan_gives_the_join(ctx, col, t):
    """Random technique per partition pair over even partitions, and the
    dichotomy's (Rs, Rs) + (R, Rl): the oracle's pairs, each pair's count the
    oracle's block join."""
    c, d = col
    ctx.set_threshold(t)
    want = O.self_join(c, t)
    parts = H.partition_even(c.n, 2000, 4)
    ij = [(i, j) for i in range(len(parts)) for j in range(i, len(parts))]
    rng = np.random.default_rng(int(t * 10))
    tune = H.Tune(bitmap_bits=128, host_threads=2)
    for trial in range(3):
        techs = rng.integers(0, 4, len(ij)) if trial else np.arange(len(ij)) % 4
        plan = [(parts[i], parts[j], int(k)) for (i, j), k in zip(ij, techs)]
        (r, s, ov), counts, ms, wall = ctx.run_partition_pairs(d, plan, cpu_threads=3, tune=tune)
        assert _rows(r, s, ov) == _want(want), (t, trial)
        assert int(counts.sum()) == want.count and wall > 0 and np.all(ms > 0)
        for (left, right, _), n in zip(plan, counts):
            assert n == O.self_join_blocks(c, t, left, right).count
        n, counts2, _, _ = ctx.run_partition_pairs(d, plan, cpu_threads=2, tune=tune, count_only=True)
        assert n == want.count and np.array_equal(counts, counts2)
    whole, small, large = H.partition_dichotomize(c.n, 1000, 30)
    for techs in ((H.TECH_CPU, H.TECH_FGSS), (H.TECH_FGSS, H.TECH_CPU), (H.TECH_CPU, H.TECH_HYBRID),
                  (H.TECH_HYBRID, H.TECH_CPU)):
        plan = [(small, small, techs[0]), (whole, large, techs[1])]
        (r, s, ov), counts, _, _ = ctx.run_partition_pairs(d, plan, cpu_threads=2, tune=tune)
        assert _rows(r, s, ov) == _want(want), techs


def test_cost_plan_end_to_end_and_overflow_keeps_result(ctx, col):
    """cost::run's plan runs to the oracle's pairs; a result larger than the
    caller's arrays stays on the context for hs_copy_last_result."""
    c, d = col
    ctx.set_threshold(0.7)
    want = _want(O.self_join(c, 0.7))
    parts = H.partition_even(c.n, 2500, 3)
    coefs = np.zeros((4, 9))
    coefs[:, 0] = [5.0, 9.0, 8.0, 0.0]            # constant costs ...
    coefs[H.TECH_CPU, 4] = 1e-3                   # ... the CPU's grows with the pair's size
    ij, tech, _, _ = ctx.cost_run(d, parts, coefs)
    assert set(tech.tolist()) <= {H.TECH_FGSS, H.TECH_CPU}
    plan = [(parts[i], parts[j], int(k)) for (i, j), k in zip(ij, tech)]
    (r, s, ov), counts, _, _ = ctx.run_partition_pairs(d, plan, cpu_threads=2)
    assert _rows(r, s, ov) == want
    assert ctx.last_result(len(want))[0].shape[0] == len(want)


def test_bad_plans(ctx, col):
    c, d = col
    ctx.set_threshold(0.8)
    with pytest.raises(H.HysetError):
        ctx.run_partition_pairs(d, [((0, 10), (5, c.n), H.TECH_FGSS)])  # past the end
    with pytest.raises(H.HysetError):
        ctx.run_partition_pairs(d, [((0, 10), (11, 20), 9)])
    with pytest.raises(H.HysetError):
        ctx.cost_run(d, [(0, 10)], np.zeros((4, 9)), enabled=(0, 0, 0, 0))
    with pytest.raises(H.HysetError):
        ctx.features_pair(d, (5, 4), (6, 7))


def test_analytic_partition_plan_on_measured_timings(ctx):
    """The analytic model fitted to measured partition-pair timings (every
    technique, bitmap at three widths) over 2, 3 and 5 even partitions of
    a 60K-set collection plans the 4-partition join through cost::run; the
    plan's pairs are the oracle's, and its measured time (median of three)
    is within 10% + 1 ms of the best single-technique plan's."""
    from hyset_amd import split as S
    c = H.generate_self(H.gen_params(60000, median=12, sigma=0.6, lo=2, hi=200, universe=60000, zipf=1.1, seed=8))
    d = ctx.upload(c)
    try:
        t = 0.75
        trains = [H.partition_even(c.n, 2500, p) for p in (2, 3, 5)]
        F, T, bt, counts, _ = S.measure_partition_pairs(ctx, d, trains, t, cpu_threads=1)
        assert F.shape == (3 + 6 + 15, 7) and T.shape == (24, 4) and set(bt) == {64, 128, 256}
        for k, parts in enumerate(trains):  # each partitioning's pair counts add up to the join
            lo = sum(len(p) * (len(p) + 1) // 2 for p in trains[:k])
            n = len(parts) * (len(parts) + 1) // 2
            assert int(counts[lo:lo + n].sum()) == O.self_join(c, t).count
        model = S.AnalyticModel().fit(F, T, bitmap_times=bt)
        parts = H.partition_even(c.n, 2500, 4)
        plan = S.partition_plan(ctx, d, parts, model, t)
        want = _want(O.self_join(c, t))
        (r, s, ov), _, _, _ = ctx.run_partition_pairs(d, plan, cpu_threads=3)
        assert _rows(r, s, ov) == want

        def wall(p):
            runs = []
            for _ in range(3):
                n, _, _, w = ctx.run_partition_pairs(d, p, cpu_threads=3, count_only=True)
                assert n == len(want)
                runs.append(w)
            return float(np.median(runs))

        fixed = {k: wall([(a, b, k) for a, b, _ in plan]) for k in range(4)}
        got = wall(plan)
        assert got <= 1.1 * min(fixed.values()) + 1.0, (got, fixed, [k for _, _, k in plan])
    finally:
        d.free()
