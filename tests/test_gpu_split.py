"""GPU tests of the cost-model layer: device features and MLP kernels
against the oracle, the learned model's gain over the analytic one, and
exact results whatever split is chosen."""
import numpy as np
import pytest

import hyset_amd as H
from hyset_amd import split as S
import oracle
from oracle import cost as OC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = H.Context(0)
    yield c
    c.close()


def _gen(n, seed, median=10, zipf=1.1, universe=None):
    return H.generate_self(H.gen_params(n, median=median, sigma=0.6, lo=2, hi=200,
                                        universe=universe or n, zipf=zipf, seed=seed))


@pytest.mark.parametrize("t,sim", [(0.8, "jaccard"), (0.6, "jaccard"), (0.75, "cosine")])
def test_features_match_oracle(ctx, t, sim):
    c = _gen(20_000, 11)
    d = ctx.upload(c)
    ctx.set_threshold(t, sim)
    for rng in (None, (0, 0), (100, 4_999), (19_000, 19_999)):
        got = ctx.features(d, rng)
        first, last = (0, c.n - 1) if rng is None else rng
        want = OC.features(c, t, first, last, sim)
        # float32 outputs of float64 device sums
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-6, err_msg=f"range {rng}")


def _net(ctx, shape=(7, 16, 12, 4), seed=3):
    return ctx.mlp(*shape, seed=seed)


def test_mlp_forward_and_gradient_match_oracle(ctx):
    shape = (7, 16, 12, 4)
    m = _net(ctx, shape)
    rng = np.random.default_rng(1)
    X = rng.normal(0, 2, (300, 7)).astype(np.float32)
    Y = rng.normal(0, 1, (300, 4)).astype(np.float32)
    mean, stdv = X.mean(0), X.std(0)
    m.set_normalization(mean, stdv)
    p = m.params()
    assert p.shape == (m.num_params,) == (16 * 7 + 16 + 12 * 16 + 12 + 4 * 12 + 4,)
    P = m.predict(X)
    Pw, _ = OC.mlp_forward(p, shape, X, mean, stdv)
    np.testing.assert_allclose(P, Pw, rtol=1e-4, atol=1e-5)
    loss, g = m.loss_grad(X, Y)
    lw, gw = OC.mlp_loss_grad(p, shape, X, Y, mean, stdv)
    assert loss == pytest.approx(lw, rel=1e-5)
    # float32 kernels vs the float64 oracle: per-element tolerance scaled by the gradient's size
    np.testing.assert_allclose(g, gw, rtol=1e-4, atol=1e-5 * np.abs(gw).max())


def test_mlp_training_reduces_loss_and_is_deterministic(ctx):
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (500, 7)).astype(np.float32)
    Y = np.stack([np.sin(3 * X[:, 0]) + X[:, 1] ** 2, X[:, 2] * X[:, 3], np.abs(X[:, 4]), X[:, 5]], 1)
    outs = []
    for _ in range(2):
        m = _net(ctx, seed=9)
        l0, _ = m.loss_grad(X, Y)
        l1 = m.train(X, Y, epochs=1500, lr=1e-2)
        assert l1 < 0.1 * l0, (l0, l1)
        outs.append(m.params())
    np.testing.assert_array_equal(outs[0], outs[1])
    # parameters after training still give the oracle's loss
    lw = OC.mlp_loss(outs[0], (7, 16, 12, 4), X, Y)
    assert lw == pytest.approx(l1, rel=1e-4)


def _synthetic_costs(F):
    """A deterministic cost landscape with the shapes measured joins show:
    device techniques with fixed launch overheads and near-linear growth,
    the CPU with none and superlinear growth in the entry count."""
    n, entries, fstd, t = F[:, 0], F[:, 5], F[:, 3], F[:, 6]
    fgss = 0.25 + 2e-5 * n + 1e-6 * entries * (1 + fstd / 50)
    bitmap = 0.15 + 2e-9 * n * n * (1.2 - t)
    hybrid = 0.9 + 4e-6 * entries * (1 + t)
    cpu = 0.002 + 4e-9 * entries ** 1.7 * (1 + fstd / 20)
    return np.stack([fgss, bitmap, hybrid, cpu], 1)


def _synthetic_features(k, seed):
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(1e3), np.log(2e5), k))
    med = rng.uniform(8, 40, k)
    t = rng.choice([0.6, 0.7, 0.8, 0.9], k)
    entries = n * np.maximum(1, med * (1 - t))
    fstd = rng.uniform(1, 60, k)
    return np.stack([n, med, 0.6 * med, fstd, n, entries, t], 1)


def test_learned_split_gains_over_analytic_on_synthetic_costs(ctx):
    Ftr, Fte = _synthetic_features(400, 1), _synthetic_features(200, 2)
    Ttr, Tte = _synthetic_costs(Ftr), _synthetic_costs(Fte)
    sp = S.Split(ctx, epochs=3000, lr=3e-3).fit(Ftr, Ttr)
    assert sp.use_learned, sp.val
    pa = sp.analytic.predict(Fte)
    pl = sp.learned.predict(Fte)
    assert S.log_error(pl, Tte) < 0.5 * S.log_error(pa, Tte)
    best = Tte.min(1).sum()
    reg_a = S.plan_cost(S.argmin_plan(pa), Tte) - best
    reg_l = S.plan_cost(S.argmin_plan(pl), Tte) - best
    assert reg_l < reg_a, (reg_l, reg_a)
    j = sp.judge(Fte, Tte)
    assert j["learned_mispredicted"] < j["analytic_mispredicted"]
    assert j["learned_right"] > j["analytic_right"]


def test_split_results_are_exact_for_any_plan(ctx):
    colls = [_gen(n, 30 + i, median=m) for i, (n, m) in enumerate([(1_000, 8), (3_000, 20), (8_000, 12),
                                                                    (2_000, 40), (12_000, 10), (5_000, 30)])]
    units = [ctx.upload(c) for c in colls]
    t = 0.75
    ctx.set_threshold(t)
    want = np.array([oracle.self_join(c, t).count for c in colls], dtype=np.uint64)
    k = len(units)
    rng = np.random.default_rng(4)
    plans = [np.full(k, H.TECH_CPU), np.full(k, H.TECH_FGSS), np.full(k, H.TECH_BITMAP),
             np.full(k, H.TECH_HYBRID), rng.integers(0, 4, k), rng.integers(0, 4, k)]
    for plan in plans:
        counts, unit_ms, wall = ctx.split_run(units, plan, cpu_threads=3)
        np.testing.assert_array_equal(counts, want, err_msg=f"plan {plan}")
        assert wall > 0 and np.all(unit_ms >= 0)


def test_measured_split_end_to_end(ctx):
    colls = [_gen(n, 50 + i, median=m, zipf=z) for i, (n, m, z) in
             enumerate([(1_000, 8, 1.0), (2_000, 30, 1.2), (4_000, 12, 0.9), (6_000, 20, 1.1), (10_000, 8, 1.3),
                        (1_500, 35, 0.8), (3_000, 15, 1.0), (8_000, 25, 1.2)])]
    units = [ctx.upload(c) for c in colls]
    F, T, counts = S.measure_units(ctx, units, 0.75)
    for c, n in zip(colls, counts):
        assert n == oracle.self_join(c, 0.75).count
    assert F.shape == (8, 7) and T.shape == (8, 4) and np.all(np.isfinite(T)) and np.all(T > 0)
    sp = S.Split(ctx, epochs=500).fit(F, T)
    plan = S.concurrent_plan(sp.predict(F), 3)
    got, _, _ = ctx.split_run(units, plan, cpu_threads=3)
    np.testing.assert_array_equal(got, counts.astype(np.uint64))
    j = sp.judge(F, T)
    assert j["units"] == 8 and j["learned_right"] + j["analytic_right"] <= j["disagree"] <= 8


def test_learned_plan_on_measured_timings(ctx):
    """The split trained on measured unit timings (no synthetic costs): on
    held-out units the learned model's concurrent plan, run by
    hs_split_run with the concurrency scales calibrated on the training
    units, takes no longer than the analytic model's plan or every unit on
    the device (median of three runs each, within 10% + 1 ms), and every
    plan's pair counts are exact."""
    rng = np.random.default_rng(21)
    specs = [(int(np.exp(rng.uniform(np.log(300), np.log(20_000)))), int(rng.integers(6, 40)),
              float(rng.uniform(0.8, 1.3))) for _ in range(48)]
    colls = [_gen(n, 200 + i, median=m, zipf=z) for i, (n, m, z) in enumerate(specs)]
    units = [ctx.upload(c) for c in colls]
    try:
        F, T, counts = S.measure_units(ctx, units, 0.75)
        te = np.arange(0, 48, 3)  # 16 held-out units
        tr = np.setdiff1d(np.arange(48), te)  # 32 training units
        sp = S.Split(ctx, epochs=1500).fit(F[tr], T[tr])
        threads = 3
        cal = S.calibrate_concurrency(ctx, [units[i] for i in tr], T[tr], threads)
        assert cal["gpu_alone"] > 0 and cal["gpu_loaded"] >= cal["gpu_alone"] and cal["cpu"] >= 1.0
        ev = [units[i] for i in te]

        def plan_of(pred):
            return S.concurrent_plan(pred, threads, cal["cpu"], cal["gpu_loaded"], cal["gpu_alone"])

        plans = {"learned": plan_of(sp.learned.predict(F[te])), "analytic": plan_of(sp.analytic.predict(F[te])),
                 "device": np.full(len(te), H.TECH_FGSS, dtype=np.int32)}
        wall = {}
        for name, plan in plans.items():
            runs = []
            for _ in range(3):
                got, _, w = ctx.split_run(ev, plan, cpu_threads=threads)
                np.testing.assert_array_equal(got, counts[te].astype(np.uint64), err_msg=name)
                runs.append(w)
            wall[name] = float(np.median(runs))
        assert wall["learned"] <= 1.1 * wall["analytic"] + 1.0, wall
        assert wall["learned"] <= 1.1 * wall["device"] + 1.0, wall
    finally:
        for u in units:
            u.free()

