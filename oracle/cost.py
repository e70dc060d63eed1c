"""CPU oracle for the cost models of the CPU/GPU split -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ as the checker of the library's feature extraction
(hs_features) and of its MLP kernels (hs_mlp_*); nothing in the product
imports it.

* features(): restates hyset::statistics::extract(partition)
  (include/hyset/statistics.hpp:163-213) in float64 numpy: the set-size
  mean and population standard deviation over sizes[startID, endID)
  (endID excluded, as thrust::transform_reduce there is called on
  [startID, endID)), size = endID - startID + 1, the population standard
  deviation over the universe of the per-token counts of the partition's
  prefix entries (index::histogram over the midprefix entries of
  collection.hpp:136-150), the entry count and the universe size.
  Pinned by hand-computed cases in tests/test_cost_cpu.py.
* features_pair(): restates statistics::extract(left, right)
  (include/hyset/statistics.hpp:216-283): a zero-filled vector of
  |left| + |right| sizes into which sizes[leftStart, leftEnd) and
  sizes[rightStart, rightEnd) are copied (each range's last set left out,
  its slot left zero), its mean and population standard deviation; the
  token-count deviation over both ranges' midprefix entries; indexSize =
  the left range's entries.  Pinned by a hand-computed case.
* partition_even(), partition_dichotomize(): restate
  include/hyset/partitioner.hpp:11-68 over blocks of consecutive sets
  (collection.hpp:147-172); pinned by hand-computed cuts and by the
  cover property (every set in exactly one partition).
* cost_run(): restates include/hyset/cost.hpp:28-58 with the linear cost
  of statistics.hpp:150-160 (|sum coeff * term + intercept|, size^2 as an
  unsigned 32-bit product when wrap_square), the first of equal costs in
  the std::map order bitmap, cpu, fgssjoin, hybrid.  Pinned by
  hand-computed costs, a tie, and the wrap of 70,000^2.
* mlp_*(): forward pass, mean-squared-error loss and reverse-mode gradient
  of the network in float64 with the library's parameter layout
  (W1 (h1 x in), b1, W2 (h2 x h1), b2, W3 (out x h2), b3; ReLU on the two
  hidden layers; inputs standardised by (x - mean) / std).  Pinned against
  central finite differences of mlp_loss in tests/test_cost_cpu.py.
"""
from __future__ import annotations

import numpy as np

from . import bounds as _bounds


def midprefix_lengths(sizes, t, sim="jaccard", mode=0):
    """midprefix(|x|) per set (include/hyset/similarity.hpp: the self-join index prefix)."""
    sizes = np.asarray(sizes, dtype=np.int64)
    out = np.empty_like(sizes)
    cache = {}
    for i, l in enumerate(sizes):
        l = int(l)
        if l not in cache:
            cache[l] = _bounds(sim, t, l, l, mode)[4]
        out[i] = cache[l]
    return out


def features(c, t, first=0, last=None, sim="jaccard", mode=0, universe=None):
    """The 7 features of hs_features for sets [first, last] of collection c."""
    last = c.n - 1 if last is None else last
    n = last - first + 1
    sizes = np.asarray(c.sizes, dtype=np.float64)
    m_last = last - 1 if n > 1 else last
    sz = sizes[first:m_last + 1]
    mean = sz.mean()
    std = np.sqrt(((sz - mean) ** 2).mean())
    U = int(universe if universe is not None else (int(np.max(c.tokens)) + 1 if len(c.tokens) else 0))
    counts = np.zeros(U, dtype=np.float64)
    plen = midprefix_lengths(c.sizes[first:last + 1], t, sim, mode)
    entries = 0
    for k, i in enumerate(range(first, last + 1)):
        st = int(c.starts[i])
        toks = np.asarray(c.tokens[st:st + int(plen[k])], dtype=np.int64)
        counts[toks] += 1
        entries += int(plen[k])
    fm = counts.mean() if U else 0.0
    fstd = np.sqrt(((counts - fm) ** 2).mean()) if U else 0.0
    return np.array([n, mean, std, fstd, U, entries, t], dtype=np.float64)


def features_pair(c, t, left, right, sim="jaccard", mode=0, universe=None):
    """The 7 features of hs_features_pair for set ranges left = (a0, a1) and
    right = (b0, b1), inclusive ends."""
    (a0, a1), (b0, b1) = left, right
    na, nb = a1 - a0 + 1, b1 - b0 + 1
    sizes = np.asarray(c.sizes, dtype=np.float64)
    vec = np.zeros(na + nb, dtype=np.float64)
    vec[:na - 1] = sizes[a0:a1]
    vec[na:na + nb - 1] = sizes[b0:b1]
    mean = vec.mean()
    std = np.sqrt(((vec - mean) ** 2).mean())
    U = int(universe if universe is not None else (int(np.max(c.tokens)) + 1 if len(c.tokens) else 0))
    counts = np.zeros(U, dtype=np.float64)
    left_entries = 0
    for lo, hi, is_left in ((a0, a1, True), (b0, b1, False)):
        plen = midprefix_lengths(c.sizes[lo:hi + 1], t, sim, mode)
        for k, i in enumerate(range(lo, hi + 1)):
            st = int(c.starts[i])
            np.add.at(counts, np.asarray(c.tokens[st:st + int(plen[k])], dtype=np.int64), 1)
            if is_left:
                left_entries += int(plen[k])
    fm = counts.mean() if U else 0.0
    fstd = np.sqrt(((counts - fm) ** 2).mean()) if U else 0.0
    return np.array([na + nb, mean, std, fstd, U, left_entries, t], dtype=np.float64)


def _blocks(n, block):
    return [(b, min(n, b + block) - 1) for b in range(0, n, block)]


def partition_even(n, block, parts):
    """partitioner::even: ceil(blocks / parts) consecutive blocks each; set-id ranges."""
    bl = _blocks(n, block)
    per = -(-len(bl) // parts)
    return [(bl[i][0], bl[min(len(bl), i + per) - 1][1]) for i in range(0, len(bl), per)]


def partition_dichotomize(n, block, split_pct):
    """partitioner::dichotomize: (whole, Rs = first ceil(split/100 * blocks) blocks, Rl); None when empty."""
    import math
    bl = _blocks(n, block)
    portion = min(len(bl), math.ceil(split_pct / 100.0 * len(bl)))
    whole = (0, n - 1) if bl else None
    small = (0, bl[portion - 1][1]) if portion > 0 else None
    large = (bl[portion][0], n - 1) if portion < len(bl) else None
    return whole, small, large


TECH_ORDER = (1, 3, 0, 2)  # bitmap, cpu, fgssjoin, hybrid as HS_TECH_* (FGSS 0, BITMAP 1, HYBRID 2, CPU 3)


def linear_cost(coef, f, bitmap, wrap_square=False):
    """coef = (intercept, set_size_std, average_set_size, token_freq, size, size_square,
    universe_size, index_size, bitmap); f = a features row."""
    size, universe, index = int(f[0]), int(f[4]), int(f[5])
    sq = (size * size) % (1 << 32) if wrap_square else size * size
    v = (coef[1] * float(np.float32(f[2])) + coef[2] * float(np.float32(f[1])) + coef[3] * float(np.float32(f[3]))
         + coef[4] * size + coef[5] * sq + coef[6] * universe + coef[7] * index + coef[8] * bitmap + coef[0])
    return abs(v)


def cost_run(feature_rows, coefs, enabled, bitmap, wrap_square=False):
    """Technique (HS_TECH_*) of least cost per row; ties to the first in TECH_ORDER."""
    out = []
    for f in feature_rows:
        best, best_c = -1, None
        for k in TECH_ORDER:
            if not enabled[k]:
                continue
            c = linear_cost(coefs[k], f, bitmap, wrap_square)
            if best < 0 or c < best_c:
                best, best_c = k, c
        out.append(best)
    return out


# ------------------------------------------------------------------- MLP

def mlp_unpack(p, n_in, h1, h2, n_out):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = []
    for rows, cols in ((h1, n_in), (h2, h1), (n_out, h2)):
        out.append(p[o:o + rows * cols].reshape(rows, cols))
        o += rows * cols
        out.append(p[o:o + rows])
        o += rows
    assert o == p.shape[0]
    return out


def mlp_forward(p, shape, X, mean=None, stdv=None):
    n_in, h1, h2, n_out = shape
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p, *shape)
    X = np.asarray(X, dtype=np.float64)
    mean = np.zeros(n_in) if mean is None else np.asarray(mean, dtype=np.float64)
    stdv = np.ones(n_in) if stdv is None else np.asarray(stdv, dtype=np.float64)
    Xs = (X - mean) / stdv
    Z1 = Xs @ W1.T + b1
    A1 = np.maximum(Z1, 0)
    Z2 = A1 @ W2.T + b2
    A2 = np.maximum(Z2, 0)
    P = A2 @ W3.T + b3
    return P, (Xs, Z1, A1, Z2, A2)


def mlp_loss(p, shape, X, Y, mean=None, stdv=None) -> float:
    P, _ = mlp_forward(p, shape, X, mean, stdv)
    return float(((P - np.asarray(Y, dtype=np.float64)) ** 2).mean())


def mlp_loss_grad(p, shape, X, Y, mean=None, stdv=None):
    """(loss, gradient) by reverse-mode differentiation, float64."""
    n_in, h1, h2, n_out = shape
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p, *shape)
    P, (Xs, Z1, A1, Z2, A2) = mlp_forward(p, shape, X, mean, stdv)
    Y = np.asarray(Y, dtype=np.float64)
    D = P - Y
    loss = float((D ** 2).mean())
    dP = 2.0 * D / D.size
    gW3 = dP.T @ A2
    gb3 = dP.sum(0)
    dZ2 = (dP @ W3) * (Z2 > 0)
    gW2 = dZ2.T @ A1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (Z1 > 0)
    gW1 = dZ1.T @ Xs
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
