"""CPU tests of the cost-model layer: the oracle pinned against hand-computed
features and finite differences, the host linear model against numpy, and
the planners against brute force."""
import itertools

import numpy as np
import pytest

import hyset_amd as H
from hyset_amd import split as S
from oracle import bounds as or_bounds
from oracle import cost as OC


def _coll(sets):
    return H.SetCollection.from_sets(sets)


def test_oracle_features_hand_computed():
    # sizes 2, 3, 3, 5; Jaccard 0.8 -> midprefix = l - ceil(2*0.8*l/1.8) + 1
    sets = [[0, 1], [0, 2, 3], [1, 2, 4], [0, 1, 2, 3, 5]]
    c = _coll(sets)
    mids = [or_bounds("jaccard", 0.8, l)[4] for l in (2, 3, 3, 5)]
    assert mids == [1, 1, 1, 1]
    f = OC.features(c, 0.8)
    # size stats over sizes[0:3] (the last set is excluded as in statistics.hpp:179-180)
    sz = np.array([2, 3, 3], dtype=float)
    assert f[0] == 4
    assert f[1] == pytest.approx(sz.mean())
    assert f[2] == pytest.approx(np.sqrt(((sz - sz.mean()) ** 2).mean()))
    # prefix tokens: 0, 0, 1, 0 over universe 6 -> counts [3, 1, 0, 0, 0, 0]
    cnt = np.array([3, 1, 0, 0, 0, 0], dtype=float)
    assert f[3] == pytest.approx(np.sqrt(((cnt - cnt.mean()) ** 2).mean()))
    assert f[4] == 6 and f[5] == 4 and f[6] == pytest.approx(0.8)
    # a sub-range and the one-set range
    g = OC.features(c, 0.8, first=1, last=2)
    assert g[0] == 2 and g[1] == 3 and g[2] == 0 and g[5] == 2
    one = OC.features(c, 0.8, first=3, last=3)
    assert one[0] == 1 and one[1] == 5 and one[2] == 0


def test_oracle_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    shape = (5, 7, 6, 3)
    n_par = 7 * 5 + 7 + 6 * 7 + 6 + 3 * 6 + 3
    p = rng.normal(0, 0.5, n_par)
    X = rng.normal(0, 1, (11, 5))
    Y = rng.normal(0, 1, (11, 3))
    mean, stdv = rng.normal(0, 0.3, 5), rng.uniform(0.5, 2, 5)
    loss, g = OC.mlp_loss_grad(p, shape, X, Y, mean, stdv)
    assert loss == pytest.approx(OC.mlp_loss(p, shape, X, Y, mean, stdv))
    eps = 1e-6
    fd = np.empty_like(p)
    for i in range(n_par):
        e = np.zeros_like(p)
        e[i] = eps
        fd[i] = (OC.mlp_loss(p + e, shape, X, Y, mean, stdv) - OC.mlp_loss(p - e, shape, X, Y, mean, stdv)) / (2 * eps)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_oracle_mlp_layout_is_the_textbook_network():
    # with identity-like weights the network reduces to relu(relu(x)) summed
    shape = (2, 2, 2, 1)
    p = np.array([1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 1, 0.5])
    P, _ = OC.mlp_forward(p, shape, np.array([[1.0, -2.0], [3.0, 4.0]]))
    np.testing.assert_allclose(P[:, 0], [1.5, 7.5])


def test_linear_fit_matches_numpy_ridge():
    rng = np.random.default_rng(7)
    X = rng.normal(0, 1, (200, 6))
    Y = X @ rng.normal(0, 1, (6, 3)) + rng.normal(0, 1, 3) + 0.01 * rng.normal(0, 1, (200, 3))
    for ridge in (0.0, 0.5):
        coef = H.linear_fit(X, Y, ridge)
        A = np.hstack([X, np.ones((200, 1))])
        R = ridge * np.eye(7)
        R[6, 6] = 0.0  # the intercept is not penalised
        want = np.linalg.solve(A.T @ A + R, A.T @ Y)
        np.testing.assert_allclose(coef, want, rtol=1e-8, atol=1e-10)
        np.testing.assert_allclose(H.linear_predict(coef, X), np.abs(A @ want), rtol=1e-9, atol=1e-12)


def test_linear_fit_rejects_singular_system():
    X = np.ones((10, 2))
    with pytest.raises(H.HysetError):
        H.linear_fit(X, np.arange(10.0), 0.0)


def test_argmin_plan_is_cost_run():
    pred = np.array([[3.0, 1.0, 2.0, 9.0], [1.0, 1.5, 2.0, 0.5]])
    np.testing.assert_array_equal(S.argmin_plan(pred), [1, 3])


def test_concurrent_plan_never_worse_than_all_device_by_prediction():
    rng = np.random.default_rng(5)
    for _ in range(20):
        pred = rng.lognormal(0, 1.5, (12, 4))
        for threads in (1, 3):
            plan = S.concurrent_plan(pred, threads)
            gbest = np.argmin(pred[:, :3], axis=1)
            assert S.plan_cost(plan, pred, threads) <= S.plan_cost(gbest, pred, threads) + 1e-9


def test_concurrent_plan_falls_back_when_sharing_does_not_pay():
    """With the measured scale of a device-only run, a mixed plan is kept
    only when its predicted makespan beats that run by the margin."""
    # four units: 10 ms on the device, 30 ms on one host thread; 2 threads
    pred = np.array([[10.0, 50, 50, 30.0]] * 4)
    mixed = S.concurrent_plan(pred, 2)
    assert (mixed == H.TECH_CPU).sum() == 1  # makespan 30 against 40 for the device alone
    # device units 2.5x slower while the host threads run: sharing no longer pays
    assert (S.concurrent_plan(pred, 2, cpu_scale=1.0, gpu_scale=2.5, gpu_alone_scale=1.0) == H.TECH_FGSS).all()
    # without that slowdown the mixed plan is kept
    np.testing.assert_array_equal(S.concurrent_plan(pred, 2, cpu_scale=1.0, gpu_scale=1.0, gpu_alone_scale=1.0),
                                  mixed)
    # the measured cost of each plan uses the scale of the run it makes
    assert S.plan_cost(np.zeros(4, dtype=np.int32), pred, 2, 1.0, 2.5, 1.0) == pytest.approx(40.0)
    assert S.plan_cost(mixed, pred, 2, 1.0, 2.5, 1.0) == pytest.approx(75.0)


def test_plan_cost_makespan_brute_force():
    T = np.array([[2.0, 9, 9, 3], [4.0, 9, 9, 1], [1.0, 9, 9, 5], [3.0, 9, 9, 2]])
    best = min(S.plan_cost(np.array(p), T, 1)
               for p in itertools.product([0, 3], repeat=4))
    plan = S.concurrent_plan(T, 1)
    assert S.plan_cost(plan, T, 1) == pytest.approx(best)
    assert S.plan_cost(np.array([0, 0, 0, 0]), T) == 10.0


# ------------------------------------------------- partition pairs (cost::run)

def test_partitioners_hand_cut_and_library_agree():
    # 25 sets in blocks of 10: [0-9] [10-19] [20-24]
    assert OC.partition_even(25, 10, 2) == [(0, 19), (20, 24)]  # ceil(3/2) = 2 blocks each
    assert OC.partition_even(25, 10, 3) == [(0, 9), (10, 19), (20, 24)]
    assert OC.partition_even(25, 10, 7) == [(0, 9), (10, 19), (20, 24)]  # fewer partitions than asked
    assert OC.partition_even(30, 10, 3) == [(0, 9), (10, 19), (20, 29)]
    assert OC.partition_dichotomize(25, 10, 50) == ((0, 24), (0, 19), (20, 24))  # ceil(1.5) = 2 blocks
    assert OC.partition_dichotomize(25, 10, 0) == ((0, 24), None, (0, 24))
    assert OC.partition_dichotomize(25, 10, 100) == ((0, 24), (0, 24), None)
    assert OC.partition_dichotomize(25, 10, 10) == ((0, 24), (0, 9), (10, 24))  # ceil(0.3) = 1 block
    rng = np.random.default_rng(5)
    for _ in range(300):
        n, block = int(rng.integers(1, 5000)), int(rng.integers(1, 700))
        parts, split = int(rng.integers(1, 40)), float(rng.uniform(0, 100))
        ev = H.partition_even(n, block, parts)
        assert ev == OC.partition_even(n, block, parts)
        # a cover of [0, n) by consecutive ranges, each a whole number of blocks but the last
        assert ev[0][0] == 0 and ev[-1][1] == n - 1 and len(ev) <= parts
        assert all(a[1] + 1 == b[0] and b[0] % block == 0 for a, b in zip(ev, ev[1:]))
        d = H.partition_dichotomize(n, block, split)
        assert d == OC.partition_dichotomize(n, block, split)
        small, large = d[1], d[2]
        assert (small[1] + 1 if small else 0) == (large[0] if large else n)
    with pytest.raises(H.HysetError):
        H.partition_even(10, 0, 2)
    with pytest.raises(H.HysetError):
        H.partition_dichotomize(10, 5, 101)


def test_oracle_features_pair_hand_computed():
    # sizes 2, 3, 3, 5, 5; Jaccard 0.8 -> midprefix 1 for every size here
    sets = [[0, 1], [0, 2, 3], [1, 2, 4], [0, 1, 2, 3, 5], [2, 3, 4, 5, 6]]
    c = _coll(sets)
    assert [or_bounds("jaccard", 0.8, l)[4] for l in (2, 3, 5)] == [1, 1, 1]
    f = OC.features_pair(c, 0.8, (0, 1), (2, 4))
    # the reference's vector: sizes[0:1] + [0] + sizes[2:4] + [0] = [2, 0, 3, 5, 0]
    v = np.array([2, 0, 3, 5, 0], dtype=float)
    assert f[0] == 5
    assert f[1] == pytest.approx(v.mean())
    assert f[2] == pytest.approx(np.sqrt(((v - v.mean()) ** 2).mean()))
    # midprefix tokens of all five sets: 0, 0, 1, 0, 2 over universe 7 -> counts [3, 1, 1, 0, 0, 0, 0]
    cnt = np.array([3, 1, 1, 0, 0, 0, 0], dtype=float)
    assert f[3] == pytest.approx(np.sqrt(((cnt - cnt.mean()) ** 2).mean()))
    assert f[4] == 7 and f[5] == 2 and f[6] == pytest.approx(0.8)  # indexSize: the left range only
    # one-set ranges: both slots zero
    g = OC.features_pair(c, 0.8, (3, 3), (4, 4))
    assert g[0] == 2 and g[1] == 0 and g[2] == 0 and g[5] == 1


def test_oracle_cost_run_hand_computed():
    f = [np.array([4, 10.0, 2.0, 3.0, 100, 50, 0.8])]  # size 4, avg 10, std 2, tok 3, U 100, index 50
    coefs = np.zeros((4, 9))
    coefs[H.TECH_FGSS] = [1, 0, 0, 0, 0, 0, 0, 0.5, 0]           # |1 + 0.5*50| = 26
    coefs[H.TECH_BITMAP] = [0, 0, 0, 0, 0, 0, 0, 0, 0.25]        # |0.25 * 64| = 16
    coefs[H.TECH_HYBRID] = [-30, 1, 1, 1, 0, 0, 0, 0, 0]         # |-30 + 2 + 10 + 3| = 15
    coefs[H.TECH_CPU] = [0, 0, 0, 0, 1, 1, 0, 0, 0]              # |4 + 16| = 20
    assert OC.linear_cost(coefs[H.TECH_HYBRID], f[0], 64) == pytest.approx(15)
    assert OC.cost_run(f, coefs, (1, 1, 1, 1), 64) == [H.TECH_HYBRID]
    assert OC.cost_run(f, coefs, (1, 1, 0, 1), 64) == [H.TECH_BITMAP]
    # a tie between cpu and fgssjoin goes to cpu (map order bitmap, cpu, fgssjoin, hybrid)
    coefs[H.TECH_CPU] = [26, 0, 0, 0, 0, 0, 0, 0, 0]
    assert OC.cost_run(f, coefs, (1, 0, 0, 1), 64) == [H.TECH_CPU]
    assert OC.cost_run(f, coefs, (1, 0, 0, 0), 64) == [H.TECH_FGSS]
    # size^2 as the reference's unsigned product: 70000^2 mod 2^32
    big = np.array([70000, 0, 0, 0, 0, 0, 0.8])
    sq = np.zeros(9)
    sq[5] = 1.0
    assert OC.linear_cost(sq, big, 0) == 70000 ** 2
    assert OC.linear_cost(sq, big, 0, wrap_square=True) == 70000 ** 2 - (1 << 32)


def test_analytic_model_recovers_reference_tables_and_round_trips():
    """Costs generated from a known coefficient table (rows HS_TECH_*,
    columns COST_TERMS) are fitted back to that table, the bitmap term from
    timings at three signature widths; predictions are the oracle's
    cost::run costs of the fitted table; the reference's JSON layout round
    trips."""
    rng = np.random.default_rng(11)
    n = 80
    F = np.stack([rng.uniform(1e3, 5e4, n), rng.uniform(5, 40, n), rng.uniform(1, 10, n), rng.uniform(0.5, 20, n),
                  rng.choice([2e4, 5e4, 1e5], n), rng.uniform(1e3, 1e5, n), np.full(n, 0.8)], axis=1)
    F = np.round(F[:, :6]).astype(np.float32).astype(np.float64).tolist()  # hs_features rows are float32
    F = np.concatenate([np.asarray(F), np.full((n, 1), 0.8)], axis=1)
    true = np.zeros((4, 9))
    true[H.TECH_FGSS] = [0.3, 0.01, 0.02, 0.05, 1e-5, 1e-10, 2e-7, 3e-6, 0]
    true[H.TECH_BITMAP] = [1.0, 0.0, 0.01, 0.0, 2e-5, 4e-9, 0.0, 0, 0.004]
    true[H.TECH_HYBRID] = [2.0, 0.03, 0.01, 0.1, 5e-5, 1e-9, 1e-7, 1e-5, 0]
    true[H.TECH_CPU] = [0.1, 0.02, 0.05, 0.02, 8e-5, 2e-9, 0.0, 2e-5, 0]

    def costs(bits):
        return np.array([[OC.linear_cost(true[k], f, bits) for k in range(4)] for f in F])

    T = costs(64)
    bt = {b: costs(b)[:, H.TECH_BITMAP] for b in (64, 128, 256)}
    m = S.AnalyticModel(ridge=1e-12).fit(F, T, bitmap_times=bt)
    tab = m.coefficients(0.8)
    np.testing.assert_allclose(tab, true, rtol=2e-3, atol=1e-5)  # normal equations on scaled terms
    for bits in (64, 256):
        want = np.array([[OC.linear_cost(tab[k], f, bits) for k in range(4)] for f in F])
        np.testing.assert_allclose(m.predict(F, bitmap_bits=bits), want, rtol=1e-9)
        np.testing.assert_allclose(m.predict(F, bitmap_bits=bits), costs(bits), rtol=1e-4)
    d = m.to_reference_json()
    assert set(d) == {"bitmap", "cpu", "fgssjoin", "hybrid"} and set(d["cpu"]) == {"0.8"}
    assert "bitmap" in d["bitmap"]["0.8"] and "index_size" not in d["bitmap"]["0.8"]
    assert "index_size" in d["fgssjoin"]["0.8"] and "bitmap" not in d["fgssjoin"]["0.8"]
    assert set(d["cpu"]["0.8"]) == {"Intercept", "set_size_std", "avg_set_size", "token_freq", "dataset_size",
                                    "dataset_size_square", "universe_size", "index_size"}
    np.testing.assert_array_equal(S.AnalyticModel.from_reference_json(d).coefficients(0.8), tab)
    # without widths the bitmap term is constant: it folds into the intercept, predictions still fit
    m1 = S.AnalyticModel(ridge=1e-12).fit(F, T)
    np.testing.assert_allclose(m1.predict(F), T, rtol=1e-4)
