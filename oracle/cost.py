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
