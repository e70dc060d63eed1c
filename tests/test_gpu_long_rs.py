"""The long-set R-S configuration at full size: R = S = 200,000 sets of
100-700 tokens over a 50,000-token Zipf(0.8) universe, 20% of S planted
near-duplicates of R sets (5-15% token edits), cosine 0.8.

The oracle's R-S join takes hours at this size, so the checks are
size-independent: two independent techniques (fgss and the bitmap filter)
return the same pair set; every returned overlap equals |r n s| computed
exactly by sparse products and meets the oracle's minoverlap; and for a
random sample of probes the returned pairs are exactly those an exhaustive
sparse product over all of S finds (completeness)."""
import numpy as np
import pytest
import scipy.sparse as sp

import hyset_amd as H
import oracle as O
from hyset_amd import workloads as W

pytestmark = pytest.mark.gpu

T = 0.8


def _csr(c):
    indptr = np.concatenate([c.starts.astype(np.int64), [c.tokens.shape[0]]])
    data = np.ones(c.tokens.shape[0], dtype=np.float32)
    # token ids run from 1 to the universe size inclusive; c.universe is max id + 1
    m = sp.csr_matrix((data, c.tokens.astype(np.int64), indptr), shape=(c.n, c.universe))
    m.check_format(full_check=True)
    return m


def _bounds(l1, l2, cache={}):
    """(minsize(l1), maxsize(l1), minoverlap(l1, l2)) from the oracle."""
    key = (int(l1), int(l2))
    if key not in cache:
        b = O.bounds("cosine", T, key[0], key[1])
        cache[key] = (b[0], b[1], b[2])
    return cache[key]


def _passes(l1, l2, overlap):
    lo, hi, mo = _bounds(l1, l2)
    return lo <= l2 <= hi and overlap >= mo


@pytest.fixture(scope="module")
def full(ctx):
    R, S = W.long_rs()
    dR, dS = ctx.upload(R), ctx.upload(S)
    ctx.set_threshold(T, "cosine")
    got = H.sort_pairs(*ctx.join(dS, dR, algo="fgss"))
    yield R, S, dR, dS, got
    dR.free()
    dS.free()


@pytest.mark.parametrize("bits", [64, 1024])
def test_long_rs_fgss_matches_bitmap(ctx, full, bits):
    R, S, dR, dS, got = full
    ctx.set_threshold(T, "cosine")
    other = H.sort_pairs(*ctx.join(dS, dR, algo="bitmap", tune=H.Tune(bitmap_bits=bits)))
    for a, b in zip(got, other):
        assert np.array_equal(a, b)
    # the planted near-duplicates are found (about 20% of S)
    assert got[0].shape[0] >= 0.19 * S.n


@pytest.fixture(scope="module")
def mats(full):
    R, S = full[0], full[1]
    return _csr(R), _csr(S)


def test_long_rs_overlaps_exact(full, mats):
    R, S, _, _, (r, s, ov) = full
    MR, MS = mats
    exact = np.asarray(MR[r].multiply(MS[s]).sum(axis=1)).ravel().astype(np.int64)
    assert np.array_equal(exact, ov.astype(np.int64))
    assert all(_passes(a, b, o) for a, b, o in zip(R.sizes[r], S.sizes[s], ov))


def test_long_rs_sampled_probes_complete(full, mats):
    R, S, _, _, (r, s, ov) = full
    MR, MS = mats
    rng = np.random.default_rng(11)
    sample = np.sort(rng.choice(R.n, 64, replace=False))
    # include probes that have planted partners
    sample = np.unique(np.concatenate([sample, r[rng.choice(r.shape[0], 16, replace=False)]]))
    prod = (MS @ MR[sample].T).toarray()  # |S| x |sample| overlaps
    ls = S.sizes.astype(np.int64)
    for k, p in enumerate(sample):
        lp = int(R.sizes[p])
        lo, hi, _ = _bounds(lp, lp)
        mos = np.array([_bounds(lp, b)[2] for b in range(ls.max() + 1)])
        want = np.flatnonzero((prod[:, k] >= mos[ls]) & (ls >= lo) & (ls <= hi))
        have = np.sort(s[r == p])
        assert np.array_equal(have, want), f"probe {p}"
