"""The benchmark configurations as generator recipes.

headline   1,000,000 sets, lognormal sizes (median 10, sigma 0.6, [2,200]),
           1M-token Zipf(1.1) universe, seed 1 (thresholds 0.6-0.9).
long_rs    R = S = 200,000 sets of 100-700 tokens, 50,000-token Zipf(0.8)
           universe, 20% of S edited copies of R sets (5-15% token edits).
batch      512 independent self-joins: n_j = round(exp(U(ln 1e3, ln 2e5))),
           median_j ~ U[8, 40], sigma 0.6, Zipf s_j ~ U[0.8, 1.3],
           universe n_j, seed 1 + j (the per-join parameters come from
           Python's random.Random(seed)).
serving    2,000,000 indexed sets (median 12, sigma 0.6, 2M-token Zipf(1.1))
           and query sets drawn with seed 2 from the same token space.
"""
from __future__ import annotations

import concurrent.futures
import math
import os
import random

from . import SIZE_LOGNORMAL, SIZE_UNIFORM, gen_params, generate_queries, generate_rs, generate_self


def headline(n=1_000_000, seed=1):
    return generate_self(gen_params(n, median=10, sigma=0.6, lo=2, hi=200, universe=1_000_000, zipf=1.1, seed=seed))


def long_rs(n=200_000, universe=50_000):
    pr = gen_params(n, size_dist=SIZE_UNIFORM, lo=100, hi=700, universe=universe, zipf=0.8, seed=1)
    ps = gen_params(n, size_dist=SIZE_UNIFORM, lo=100, hi=700, universe=universe, zipf=0.8, seed=2)
    return generate_rs(pr, ps, dup_frac=0.2, edit_lo=0.05, edit_hi=0.15)


def batch_params(k=512, seed=1, n_lo=1_000, n_hi=200_000):
    rnd = random.Random(seed)
    out = []
    for j in range(k):
        n = int(round(math.exp(rnd.uniform(math.log(n_lo), math.log(n_hi)))))
        med = rnd.uniform(8, 40)
        s = rnd.uniform(0.8, 1.3)
        out.append(gen_params(n, median=med, sigma=0.6, lo=2, hi=400, universe=n, zipf=s, seed=seed + j))
    return out


def batch(k=512, seed=1, n_lo=1_000, n_hi=200_000, threads=None):
    params = batch_params(k, seed, n_lo, n_hi)
    threads = threads or min(16, len(os.sched_getaffinity(0)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(generate_self, params))


def serving(n_index=2_000_000, n_queries=200_000):
    pi = gen_params(n_index, median=12, sigma=0.6, lo=2, hi=200, universe=2_000_000, zipf=1.1, seed=1)
    pq = gen_params(n_queries, median=12, sigma=0.6, lo=2, hi=200, universe=2_000_000, zipf=1.1, seed=2)
    return generate_queries(pi, pq)
