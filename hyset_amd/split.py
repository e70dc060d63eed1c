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
  technique and threshold, on each technique's terms of the reference's
  coefficient file (setSizeStd, averageSetSize, tokenFreqDistribution,
  size, size^2, universeSize, and indexSize -- or, for bitmap, the
  signature width); it reads and writes that file's JSON layout, and its
  tables drive hs_cost_run (cost::run over partition pairs) as they are;
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

from . import TECH_BITMAP, TECH_CPU, TECH_FGSS, TECH_HYBRID, TECHNIQUES, Context, linear_fit

N_TECH = len(TECHNIQUES)
GPU_TECHS = (TECH_FGSS, TECH_BITMAP, TECH_HYBRID)


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

TECH_NAMES = {TECH_BITMAP: "bitmap", TECH_CPU: "cpu", TECH_FGSS: "fgssjoin", TECH_HYBRID: "hybrid"}
# the reference's JSON keys per coefficient (src/cost_model.cu:306-317), in COST_TERMS order
JSON_KEYS = ("Intercept", "set_size_std", "avg_set_size", "token_freq", "dataset_size", "dataset_size_square",
             "universe_size", "index_size", "bitmap")


def technique_terms(F: np.ndarray, tech: int, bitmap_bits) -> np.ndarray:
    """The terms of one technique's linear cost, in COST_TERMS order after
    the intercept: setSizeStd, averageSetSize, tokenFreqDistribution, size,
    size^2, universeSize, and indexSize for every technique but bitmap,
    bitmap (the signature width) for bitmap -- read_coefficients gives
    bitmap no index_size and the others no bitmap (cost_model.cu:314-315)."""
    F = np.atleast_2d(np.asarray(F, dtype=np.float64))
    size = F[:, 0]
    last = np.broadcast_to(np.asarray(bitmap_bits, dtype=np.float64), size.shape) if tech == TECH_BITMAP else F[:, 5]
    return np.stack([F[:, 2], F[:, 1], F[:, 3], size, size * size, F[:, 4], last], axis=1)


class AnalyticModel:
    """The reference's linear cost per technique and threshold
    (statistics.hpp:132-161; one table per threshold key as
    cost_model.cu:296-323), fitted by ridge least squares (hs_linear_fit)
    on column-scaled terms and stored unscaled in the reference's layout,
    so it evaluates as cost::run does and exports to the reference's
    coefficient file."""

    def __init__(self, ridge: float = 1e-6, bitmap_bits: int = 64):
        self.ridge = ridge
        self.bitmap_bits = bitmap_bits
        self.table: dict[str, np.ndarray] = {}  # threshold key -> (4, 9) in COST_TERMS order

    @staticmethod
    def _fit_one(X, y, ridge):
        # a term constant over the samples (the universe of one collection's
        # partitions, the width of one bitmap) is the intercept's: weight 0
        live = np.ptp(X, axis=0) > 1e-12 * np.maximum(np.abs(X).max(axis=0), 1e-300)
        w = np.zeros(X.shape[1])
        if not live.any():
            return w, float(np.mean(y))
        Xl = X[:, live]
        scale = np.maximum(np.abs(Xl).max(axis=0), 1e-12)
        coef = linear_fit(Xl / scale, np.asarray(y, dtype=np.float64), ridge)  # (d + 1, 1): terms, intercept
        w[live] = coef[:-1, 0] / scale
        return w, float(coef[-1, 0])

    def fit(self, F, T, bitmap_times: dict | None = None) -> "AnalyticModel":
        """F (n, 7) features, T (n, 4) times per HS_TECH_*.  bitmap_times
        {bits: (n,) times} gives the bitmap technique at several signature
        widths (its bitmap term); without it T[:, bitmap] is at
        self.bitmap_bits and the term is constant."""
        F = np.atleast_2d(F)
        T = np.asarray(T, dtype=np.float64)
        keys = np.array([threshold_key(t) for t in F[:, 6]])
        bt = bitmap_times if bitmap_times else {self.bitmap_bits: T[:, TECH_BITMAP]}
        for k in np.unique(keys):
            m = keys == k
            tab = np.zeros((4, 9))
            for tech in range(N_TECH):
                if tech == TECH_BITMAP:
                    X = np.concatenate([technique_terms(F[m], tech, b) for b in bt])
                    y = np.concatenate([np.asarray(v, dtype=np.float64)[m] for v in bt.values()])
                else:
                    X, y = technique_terms(F[m], tech, 0), T[m, tech]
                w, b0 = self._fit_one(X, y, self.ridge)
                tab[tech, 0] = b0
                tab[tech, 1:7] = w[:6]
                tab[tech, 8 if tech == TECH_BITMAP else 7] = w[6]
            self.table[k] = tab
        return self

    def coefficients(self, t: float) -> np.ndarray:
        """(4, 9) table of threshold t (the nearest fitted one), rows HS_TECH_*."""
        k = threshold_key(t)
        if k not in self.table:
            k = min(self.table, key=lambda c: abs(float(c) - float(k)))
        return self.table[k]

    def predict(self, F, bitmap_bits: int | None = None) -> np.ndarray:
        F = np.atleast_2d(F)
        bits = self.bitmap_bits if bitmap_bits is None else bitmap_bits
        out = np.empty((F.shape[0], N_TECH))
        for i, f in enumerate(F):
            tab = self.coefficients(float(f[6]))
            for tech in range(N_TECH):
                x = technique_terms(f[None, :], tech, bits)[0]
                w = np.concatenate([tab[tech, 1:7], tab[tech, 8:9] if tech == TECH_BITMAP else tab[tech, 7:8]])
                out[i, tech] = abs(tab[tech, 0] + float(w @ x))  # statistics.hpp:150-160 takes |.|
        return out

    def to_reference_json(self) -> dict:
        """The coefficient file of src/cost_model.cu:296-323:
        {technique: {threshold key: {Intercept, set_size_std, ...}}}."""
        out = {TECH_NAMES[t]: {} for t in range(N_TECH)}
        for k, tab in self.table.items():
            for tech in range(N_TECH):
                row = {JSON_KEYS[j]: float(tab[tech, j]) for j in range(7)}
                if tech == TECH_BITMAP:
                    row["bitmap"] = float(tab[tech, 8])
                else:
                    row["index_size"] = float(tab[tech, 7])
                out[TECH_NAMES[tech]][k] = row
        return out

    @classmethod
    def from_reference_json(cls, d: dict, bitmap_bits: int = 64) -> "AnalyticModel":
        m = cls(bitmap_bits=bitmap_bits)
        keys = set()
        for tech in range(N_TECH):
            keys |= set(d[TECH_NAMES[tech]])
        for k in keys:
            tab = np.zeros((4, 9))
            for tech in range(N_TECH):
                row = d[TECH_NAMES[tech]][k]
                for j in range(7):
                    tab[tech, j] = row[JSON_KEYS[j]]
                tab[tech, 8 if tech == TECH_BITMAP else 7] = row["bitmap" if tech == TECH_BITMAP else "index_size"]
            m.table[k] = tab
        return m


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


# ------------------------------------------------------- partition pairs

def measure_partition_pairs(ctx: Context, coll, partitionings, t: float, cpu_threads: int = 1,
                            bitmap_widths=(64, 128, 256), host_threads: int = 1):
    """Every partition pair (i <= j) of each partitioning (a list of set-id
    ranges of coll) at threshold t: its features (hs_cost_run's
    hs_features / hs_features_pair) and its time with every technique
    (hs_run_partition_pairs, all pairs on one technique: device pairs one
    after another, CPU pairs on cpu_threads threads -- 1 gives one-at-a-time
    times).  Returns (features (P, 7), times (P, 4) with bitmap at
    bitmap_widths[0], {width: (P,) bitmap times}, counts (P,), pairs
    [(left, right)]) and asserts every technique found each pair's count."""
    from . import Tune
    ctx.set_threshold(t)
    feats, times, bt, counts, pairs = [], [], {b: [] for b in bitmap_widths}, [], []
    zero = np.zeros((4, 9))
    for parts in partitionings:
        ij, _, f, _ = ctx.cost_run(coll, parts, zero)
        plan_pairs = [(parts[i], parts[j]) for i, j in ij]
        tt = np.zeros((len(ij), N_TECH))
        seen = None
        for tech in range(N_TECH):
            widths = bitmap_widths if tech == TECH_BITMAP else (0,)
            for b in widths:
                tune = Tune(bitmap_bits=b or 64, host_threads=host_threads)
                n, c, ms, _ = ctx.run_partition_pairs(coll, [(l, r, tech) for l, r in plan_pairs],
                                                      cpu_threads=cpu_threads, tune=tune, count_only=True)
                assert seen is None or np.array_equal(c, seen), f"technique {tech} disagrees on pair counts"
                seen = c
                if tech == TECH_BITMAP:
                    bt[b].append(ms)
                    if b == bitmap_widths[0]:
                        tt[:, tech] = ms
                else:
                    tt[:, tech] = ms
        feats.append(f.astype(np.float64))
        times.append(tt)
        counts.append(seen)
        pairs.extend(plan_pairs)
    return (np.concatenate(feats), np.concatenate(times), {b: np.concatenate(v) for b, v in bt.items()},
            np.concatenate(counts), pairs)


def partition_plan(ctx: Context, coll, parts, model: AnalyticModel, t: float, bitmap_bits: int = 64,
                   enabled=(1, 1, 1, 1)):
    """cost::run with the model's table for t (hs_cost_run): the plan
    [(left, right, technique)] over the partition pairs of parts."""
    ctx.set_threshold(t)
    ij, tech, _, _ = ctx.cost_run(coll, parts, model.coefficients(t), enabled=enabled, bitmap_bits=bitmap_bits)
    return [(parts[i], parts[j], int(k)) for (i, j), k in zip(ij, tech)]


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
           "plan_cost", "summarize", "log_error", "technique_terms", "learned_inputs", "default_threads",
           "TECH_NAMES", "JSON_KEYS", "measure_partition_pairs", "partition_plan"]
