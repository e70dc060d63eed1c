"""The CPU/GPU split of a join workload: analytic and learned cost models.

Reference: hyset::cost::run (include/hyset/cost.hpp:28-58) gives every unit
of work (a partition pair there, one self-join here) the technique in
{fgss, bitmap, hybrid, cpu} whose modelled cost is lowest.  The modelled
cost is |sum coeff * term + intercept| over the partition features
(include/hyset/statistics.hpp:132-161) with one coefficient table per
technique and threshold (src/cost_model.cu:296-323).  The reference ships
no coefficients, so:

* AnalyticModel fits those linear costs by ridge least squares (the
  library's hs_linear_fit) to timings this library measures, one table per
  technique and threshold, on the reference's eight terms (setSizeStd,
  averageSetSize, tokenFreqDistribution, size, size^2, universeSize,
  indexSize, bitmap);
* LearnedModel trains the library's MLP (hs_mlp_train: device forward and
  backward kernels, Adam) on the same timings to predict log-time per
  technique from log-features and the threshold, and predicts with the
  device inference kernel.  The analytic model stays as its fallback (when
  its held-out error is worse) and as the judge of its mispredictions.

Both models feed the same planner, so a comparison measures the models and
nothing else.  Whatever a planner chooses, every unit's pairs are those of
its own exact join: hs_split_run runs each unit with an exact technique.
"""
from __future__ import annotations

import heapq
import time

import numpy as np

from . import TECH_BITMAP, TECH_CPU, TECH_FGSS, TECH_HYBRID, TECHNIQUES, Context, linear_fit, linear_predict

N_TECH = len(TECHNIQUES)
GPU_TECHS = (TECH_FGSS, TECH_BITMAP, TECH_HYBRID)


def reference_terms(F: np.ndarray, bitmap_bits: int = 64) -> np.ndarray:
    """The eight terms of statistics.hpp:150-160 from hs_features rows."""
    F = np.atleast_2d(np.asarray(F, dtype=np.float64))
    size = F[:, 0]
    return np.stack([F[:, 2], F[:, 1], F[:, 3], size, size * size, F[:, 4], F[:, 5],
                     np.full_like(size, float(bitmap_bits))], axis=1)


def learned_inputs(F: np.ndarray) -> np.ndarray:
    """Log-scaled features plus the threshold: the network's input."""
    F = np.atleast_2d(np.asarray(F, dtype=np.float64))
    return np.concatenate([np.log1p(F[:, :6]), F[:, 6:7]], axis=1)


def threshold_key(t: float) -> str:
    return f"{t:.2g}"  # the std::setprecision(2) key of cost_model.cu:298-301


# ---------------------------------------------------------------- timing

def measure_units(ctx: Context, units, t: float, reps: int = 1, cpu_threads: int = 1, bitmap_bits: int = 64):
    """Per unit: features and the measured time (ms) of each technique.

    units are DeviceCollections; each is self-joined at threshold t with
    every technique (device ones through hs_join, the CPU one through
    hs_join_host), reps times, keeping the fastest.  Returns
    (features (n, 7) float64, times (n, 4) float64, counts (n,) int64) and
    asserts that every technique found the same number of pairs."""
    from . import Tune
    ctx.set_threshold(t)
    feats = np.stack([ctx.features(u).astype(np.float64) for u in units])
    times = np.full((len(units), N_TECH), np.inf)
    counts = np.zeros(len(units), dtype=np.int64)
    tune = Tune(bitmap_bits=bitmap_bits)
    for i, u in enumerate(units):
        seen = set()
        for tech in range(N_TECH):
            for _ in range(reps):
                a = time.perf_counter()
                if tech == TECH_CPU:
                    c = len(ctx.join_host(u, threads=cpu_threads)[0])
                else:
                    c = ctx.join(u, algo=TECHNIQUES[tech], tune=tune, count_only=True)
                times[i, tech] = min(times[i, tech], (time.perf_counter() - a) * 1e3)
                seen.add(c)
        assert len(seen) == 1, f"unit {i}: techniques disagree on the pair count {seen}"
        counts[i] = seen.pop()
    return feats, times, counts


# ---------------------------------------------------------------- models

class AnalyticModel:
    """Linear cost per technique and threshold (statistics.hpp:132-161)."""

    def __init__(self, ridge: float = 1e-6, bitmap_bits: int = 64):
        self.ridge = ridge
        self.bitmap_bits = bitmap_bits
        self.coef: dict[str, np.ndarray] = {}
        self.scale: np.ndarray | None = None

    def _terms(self, F):
        X = reference_terms(F, self.bitmap_bits)
        # the bitmap term is constant for one width: it folds into the intercept
        return X[:, :7] / self.scale

    def fit(self, F, T) -> "AnalyticModel":
        F = np.atleast_2d(F)
        X = reference_terms(F, self.bitmap_bits)[:, :7]
        self.scale = np.maximum(np.abs(X).max(axis=0), 1e-12)
        keys = np.array([threshold_key(t) for t in F[:, 6]])
        for k in np.unique(keys):
            m = keys == k
            self.coef[k] = linear_fit(self._terms(F[m]), np.asarray(T, dtype=np.float64)[m], self.ridge)
        return self

    def predict(self, F) -> np.ndarray:
        F = np.atleast_2d(F)
        out = np.empty((F.shape[0], N_TECH))
        keys = np.array([threshold_key(t) for t in F[:, 6]])
        for k in np.unique(keys):
            m = keys == k
            coef = self.coef.get(k)
            if coef is None:  # nearest fitted threshold
                coef = self.coef[min(self.coef, key=lambda c: abs(float(c) - float(k)))]
            out[m] = linear_predict(coef, self._terms(F[m]))
        return out


class LearnedModel:
    """The MLP cost model: log-time per technique from log-features."""

    def __init__(self, ctx: Context, hidden=(32, 32), seed: int = 1, epochs: int = 3000, lr: float = 3e-3):
        self.ctx = ctx
        self.hidden, self.seed = hidden, seed
        self.net = None
        self.epochs, self.lr = epochs, lr
        self.train_loss = None

    def fit(self, F, T) -> "LearnedModel":
        """Trains a freshly initialised network (same seed: same result)."""
        if self.net is not None:
            self.net.close()
        self.net = self.ctx.mlp(7, self.hidden[0], self.hidden[1], N_TECH, self.seed)
        X = learned_inputs(F)
        Y = np.log(np.maximum(np.asarray(T, dtype=np.float64), 1e-3))
        self.net.set_normalization(X.mean(0), X.std(0) + 1e-6)
        self.train_loss = self.net.train(X.astype(np.float32), Y.astype(np.float32), self.epochs, self.lr)
        return self

    def predict(self, F) -> np.ndarray:
        return np.exp(self.net.predict(learned_inputs(F).astype(np.float32)).astype(np.float64))


def log_error(pred: np.ndarray, T: np.ndarray) -> float:
    """Mean squared error of log-times (the yardstick both models are held to)."""
    return float(np.mean((np.log(np.maximum(pred, 1e-3)) - np.log(np.maximum(T, 1e-3))) ** 2))


# --------------------------------------------------------------- planning

def argmin_plan(pred: np.ndarray) -> np.ndarray:
    """cost::run (cost.hpp:28-58): each unit takes its cheapest technique."""
    return np.argmin(pred, axis=1).astype(np.int32)


def concurrent_plan(pred: np.ndarray, cpu_threads: int, cpu_scale: float = 1.0, gpu_scale: float = 1.0,
                    gpu_alone_scale: float | None = None, margin: float = 0.05) -> np.ndarray:
    """Longest-first list scheduling of units onto one device queue and
    cpu_threads host threads, by predicted time: each unit goes where it
    would finish first (the device takes its cheapest device technique).
    Runs the split the way hs_split_run executes it: CPU units concurrently
    with the device units.  The scales map the predictions (times of one
    unit at a time) to the concurrent run (calibrate_concurrency): cpu_scale
    for a CPU unit while the other host threads are busy, gpu_scale for a
    device unit while they are.  With gpu_alone_scale (a device unit in a
    run of device units alone) the plan falls back to every unit on its
    device technique unless its predicted makespan beats that run by the
    margin: sharing must pay for itself."""
    pred = np.asarray(pred, dtype=np.float64)
    gtech = np.array(GPU_TECHS)[np.argmin(pred[:, list(GPU_TECHS)], axis=1)]
    g = pred[np.arange(len(pred)), gtech]
    gl = g * gpu_scale
    c = pred[:, TECH_CPU] * cpu_scale
    plan = gtech.astype(np.int32).copy()
    gpu_load = 0.0
    cpu = [0.0] * max(1, cpu_threads)
    for i in np.argsort(-np.minimum(gl, c), kind="stable"):
        if cpu_threads > 0 and cpu[0] + c[i] < gpu_load + gl[i]:
            plan[i] = TECH_CPU
            heapq.heapreplace(cpu, cpu[0] + c[i])
        else:
            gpu_load += gl[i]
    if gpu_alone_scale is not None and (plan == TECH_CPU).any():
        alone = float(g.sum() * gpu_alone_scale)
        if max(gpu_load, max(cpu)) > alone * (1.0 - margin):
            return gtech.astype(np.int32)
    return plan


def plan_cost(plan: np.ndarray, T: np.ndarray, cpu_threads: int = 0, cpu_scale: float = 1.0,
              gpu_scale: float = 1.0, gpu_alone_scale: float = 1.0) -> float:
    """Measured cost of a plan: the sum of its units' times (sequential, as
    cost_model.cu runs them) or, with cpu_threads > 0, the makespan of the
    device queue against the host threads (list-scheduled longest first,
    CPU times stretched by cpu_scale, device times by gpu_scale when the
    plan has CPU units and by gpu_alone_scale when it has none)."""
    t = T[np.arange(len(plan)), plan]
    if cpu_threads <= 0:
        return float(t.sum())
    on_cpu = plan == TECH_CPU
    gpu = float(t[~on_cpu].sum()) * (gpu_scale if on_cpu.any() else gpu_alone_scale)
    cpu = [0.0] * cpu_threads
    for x in sorted(t[on_cpu] * cpu_scale, reverse=True):
        heapq.heapreplace(cpu, cpu[0] + x)
    return max(gpu, max(cpu))


def calibrate_cpu_scale(ctx: Context, units, T: np.ndarray, cpu_threads: int) -> float:
    """Median slowdown of CPU units run all at once on cpu_threads threads
    (hs_split_run) against their one-at-a-time times T[:, cpu]."""
    plan = np.full(len(units), TECH_CPU, dtype=np.int32)
    _, unit_ms, _ = ctx.split_run(units, plan, cpu_threads=cpu_threads)
    m = T[:, TECH_CPU] >= 1.0  # long enough to time
    if not np.any(m):
        return 1.0
    return float(max(1.0, np.median(unit_ms[m] / T[m, TECH_CPU])))


def calibrate_concurrency(ctx: Context, units, T: np.ndarray, cpu_threads: int) -> dict:
    """The scale factors of concurrent_plan, measured with hs_split_run on
    units whose one-at-a-time times T are known: 'gpu_alone', device units
    (each on its fastest device technique) run back to back with no CPU
    units; 'gpu_loaded' and 'cpu', the device and the CPU units of a mixed
    plan (the one concurrent_plan makes of T without interference) run
    together -- sums of measured over one-at-a-time times for the device
    queue, the median ratio for CPU units of at least 0.5 ms."""
    T = np.asarray(T, dtype=np.float64)
    r = np.arange(len(units))
    dev = np.array(GPU_TECHS)[np.argmin(T[:, list(GPU_TECHS)], axis=1)].astype(np.int32)
    _, ms, _ = ctx.split_run(units, dev, cpu_threads=cpu_threads)
    gpu_alone = float(ms.sum() / max(T[r, dev].sum(), 1e-9))
    mixed = concurrent_plan(T, cpu_threads, 1.0, gpu_alone)
    out = {"gpu_alone": gpu_alone, "gpu_loaded": gpu_alone, "cpu": 1.0}
    on_cpu = mixed == TECH_CPU
    if on_cpu.any() and (~on_cpu).any():
        _, ms, _ = ctx.split_run(units, mixed, cpu_threads=cpu_threads)
        out["gpu_loaded"] = float(max(ms[~on_cpu].sum() / max(T[r[~on_cpu], mixed[~on_cpu]].sum(), 1e-9),
                                      gpu_alone))
        m = on_cpu & (T[:, TECH_CPU] >= 0.5)
        if m.any():
            out["cpu"] = float(max(1.0, np.median(ms[m] / T[m, TECH_CPU])))
    return out


class Split:
    """Learned split with the analytic model as fallback and judge."""

    def __init__(self, ctx: Context, hidden=(32, 32), seed: int = 1, epochs: int = 3000, lr: float = 3e-3,
                 bitmap_bits: int = 64):
        self.analytic = AnalyticModel(bitmap_bits=bitmap_bits)
        self.learned = LearnedModel(ctx, hidden, seed, epochs, lr)
        self.use_learned = True
        self.val = {}

    def fit(self, F, T, holdout: float = 0.2, seed: int = 0) -> "Split":
        F, T = np.atleast_2d(F), np.asarray(T, dtype=np.float64)
        n = F.shape[0]
        idx = np.random.default_rng(seed).permutation(n)
        k = max(1, int(round(n * holdout))) if n >= 5 else 0
        va, tr = idx[:k], idx[k:]
        if k:
            a = AnalyticModel(bitmap_bits=self.analytic.bitmap_bits).fit(F[tr], T[tr])
            self.learned.fit(F[tr], T[tr])
            self.val = {"analytic": log_error(a.predict(F[va]), T[va]),
                        "learned": log_error(self.learned.predict(F[va]), T[va])}
            # the fallback: keep the analytic model when the network does worse on held-out units
            self.use_learned = bool(np.isfinite(self.val["learned"]) and self.val["learned"] <= self.val["analytic"])
        self.analytic.fit(F, T)
        self.learned.fit(F, T)
        return self

    def predict(self, F) -> np.ndarray:
        if self.use_learned:
            p = self.learned.predict(F)
            if np.all(np.isfinite(p)):
                return p
        return self.analytic.predict(F)

    def judge(self, F, T) -> dict:
        """Where the learned and analytic argmin choices differ, which one the
        measured times favour (the analytic model as judge of mispredictions)."""
        pl = argmin_plan(self.learned.predict(F))
        pa = argmin_plan(self.analytic.predict(F))
        best = np.argmin(T, axis=1)
        r = np.arange(len(pl))
        diff = pl != pa
        return {"units": int(len(pl)), "disagree": int(diff.sum()),
                "learned_right": int((diff & (T[r, pl] < T[r, pa])).sum()),
                "analytic_right": int((diff & (T[r, pa] < T[r, pl])).sum()),
                "learned_mispredicted": int((pl != best).sum()),
                "analytic_mispredicted": int((pa != best).sum())}


def summarize(plans: dict, T: np.ndarray, cpu_threads: int, cpu_scale: float = 1.0, gpu_scale: float = 1.0,
              gpu_alone_scale: float = 1.0) -> dict:
    """Measured cost of each plan against the best per-unit choice."""
    best = float(T.min(axis=1).sum())
    out = {}
    for name, plan in plans.items():
        seq = plan_cost(plan, T)
        out[name] = {"sum_ms": round(seq, 3), "regret_ms": round(seq - best, 3),
                     "makespan_ms": round(plan_cost(plan, T, cpu_threads, cpu_scale, gpu_scale, gpu_alone_scale),
                                          3),
                     "share": {TECHNIQUES[k]: int((plan == k).sum()) for k in range(N_TECH)}}
    return out


def default_threads() -> int:
    import os
    return max(1, min(14, len(os.sched_getaffinity(0)) - 2))


__all__ = ["AnalyticModel", "LearnedModel", "Split", "measure_units", "argmin_plan", "concurrent_plan",
           "calibrate_cpu_scale", "calibrate_concurrency",
           "plan_cost", "summarize", "log_error", "reference_terms", "learned_inputs", "default_threads"]
