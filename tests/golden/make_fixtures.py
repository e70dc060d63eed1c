"""Writes the golden fixtures under tests/golden/.

* config1.sha256 -- digest of the 10,000-set configuration (sizes uniform in
  [5, 50], 10,000-token Zipf(1.0) universe, seed 1) in the reference's text
  format, so generator drift is caught;
* config1_t{t}.pairs -- its sorted "r s overlap" result at several Jaccard
  thresholds, computed by the oracle's brute force and AllPairs, which must
  agree before anything is written;
* neardup.sha256 / neardup_t{t}.pairs -- 4,000 sets, half of them edited
  copies of the others, and its results (brute force == AllPairs);
* special.txt / special_t{t}_{sim}.pairs -- a hand-made collection (identical
  sets, disjoint sets, size-1 and size-2 sets, token 0, sets at the t = 0.8
  and t = 0.9 double-rounding sums) with its result computed by a pure-Python
  brute force that shares no code with the oracle.

The reference holds no fixtures of its own; these come from this project's
generator, the oracle and the independent Python brute force.
Run:  python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import hashlib
import math
import sys
import tempfile
from fractions import Fraction
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import numpy as np  # noqa: E402

import hyset_amd as H  # noqa: E402
import oracle as O  # noqa: E402

CONFIG1 = dict(size_dist=H.SIZE_UNIFORM, lo=5, hi=50, universe=10_000, zipf=1.0, seed=1)
CONFIG1_T = [0.3, 0.4, 0.5, 0.8]
NEARDUP_T = [0.6, 0.7, 0.8, 0.9]


def neardup():
    """4,000 sets: 2,000 random sets plus 2,000 edited copies (5-15% token
    edits) of random ones, re-sorted by (size, tokens) -- many pairs at high t."""
    r, s = H.generate_rs(H.gen_params(2000, size_dist=H.SIZE_UNIFORM, lo=5, hi=40, universe=3000, zipf=0.9,
                                      seed=11),
                         H.gen_params(2000, size_dist=H.SIZE_UNIFORM, lo=5, hi=40, universe=3000, zipf=0.9,
                                      seed=12), dup_frac=1.0, edit_lo=0.05, edit_hi=0.15)
    sets = [list(r.set(i)) for i in range(r.n)] + [list(s.set(i)) for i in range(s.n)]
    sets.sort(key=lambda x: (len(x), x))
    return H.SetCollection.from_sets(sets)


def config1():
    return H.generate_self(H.gen_params(10_000, **CONFIG1))


def collection_digest(c) -> str:
    with tempfile.NamedTemporaryFile(suffix=".txt") as f:
        H.write_collection_text(f.name, c)
        return hashlib.sha256(Path(f.name).read_bytes()).hexdigest()


def write_pairs(path: Path, r, s, ov) -> None:
    with open(path, "w") as f:
        for a, b, o in zip(r, s, ov):
            f.write(f"{a} {b} {o}\n")


def read_pairs(path: Path):
    rows = [tuple(int(x) for x in line.split()) for line in open(path) if line.strip()]
    return rows


# ------------------------------------------------------ special collection

SPECIAL_SETS = [
    [0], [0], [7],                                   # size 1, token 0, identical pair
    [1, 2], [1, 2], [1, 3],                          # size 2
    [0, 5, 9], [4, 5, 9],
    [10, 11, 12, 13], [20, 21, 22, 23],              # disjoint
    # 19 + 19 tokens sharing 18 (sum 38): J = 0.9 exactly; the double bound at t = 0.9 asks for 19
    list(range(100, 118)) + [500], list(range(100, 119)),
    list(range(200, 210)), list(range(200, 209)) + [300],   # J = 9/11
    # 32 + 31 tokens sharing 28 (sum 63): J = 0.8 exactly; the double bound at t = 0.8 asks for 29
    list(range(1000, 1032)), list(range(1000, 1028)) + [2001, 2002, 2003],
    # 63 + 63 tokens sharing 56 (sum 126): J = 0.8 exactly; the double bound asks for 57
    list(range(4000, 4063)), list(range(4000, 4056)) + list(range(4900, 4907)),
    list(range(3000, 3028)), list(range(3000, 3028)), list(range(3000, 3027)) + [3500],
]


def special_collection():
    sets = sorted((sorted(s) for s in SPECIAL_SETS), key=lambda s: (len(s), s))
    return H.SetCollection.from_sets(sets), sets


def py_bounds(sim, t, l1, l2):
    """Independent restatement: Jaccard in IEEE double as the reference
    writes it, cosine and Dice with exact fractions."""
    f = Fraction(str(t))
    if sim == "jaccard":
        mins = int(math.ceil(t * l1))
        maxs = int(l1 / t)
        mo = int(math.ceil((l1 + l2) * t / (1 + t)))
    elif sim == "dice":
        mins = math.ceil(f * l1 / (2 - f))
        maxs = math.floor((2 - f) * l1 / f)
        mo = math.ceil(f * (l1 + l2) / 2)
    else:
        mins = math.ceil(f * f * l1)
        maxs = math.floor(l1 / (f * f))
        mo = 0
        while Fraction(mo * mo) < f * f * l1 * l2:
            mo += 1
    return mins, maxs, min(l1, l2, mo)


def py_bruteforce(sets, sim, t):
    out = []
    for r in range(len(sets)):
        for s in range(r):
            ov = len(set(sets[r]) & set(sets[s]))
            mins, maxs, mo = py_bounds(sim, t, len(sets[r]), len(sets[s]))
            if mins <= len(sets[s]) <= maxs and ov >= mo:
                out.append((r, s, ov))
    return out


SPECIAL_CASES = [("jaccard", 0.8), ("jaccard", 0.9), ("jaccard", 0.5), ("cosine", 0.8), ("dice", 0.8)]


def main() -> None:
    c = config1()
    (HERE / "config1.sha256").write_text(collection_digest(c) + "\n")
    for t in CONFIG1_T:
        bf = O.bruteforce_self(c, t)
        ap = O.self_join(c, t)
        assert np.array_equal(bf.r, ap.r) and np.array_equal(bf.s, ap.s) and np.array_equal(bf.overlap, ap.overlap)
        write_pairs(HERE / f"config1_t{t}.pairs", bf.r, bf.s, bf.overlap)
        print(f"config1 t={t}: {bf.count} pairs")
    nd = neardup()
    (HERE / "neardup.sha256").write_text(collection_digest(nd) + "\n")
    for t in NEARDUP_T:
        bf = O.bruteforce_self(nd, t)
        ap = O.self_join(nd, t)
        assert np.array_equal(bf.r, ap.r) and np.array_equal(bf.s, ap.s) and np.array_equal(bf.overlap, ap.overlap)
        write_pairs(HERE / f"neardup_t{t}.pairs", bf.r, bf.s, bf.overlap)
        print(f"neardup t={t}: {bf.count} pairs")
    coll, sets = special_collection()
    H.write_collection_text(HERE / "special.txt", coll)
    for sim, t in SPECIAL_CASES:
        rows = py_bruteforce(sets, sim, t)
        write_pairs(HERE / f"special_t{t}_{sim}.pairs", *zip(*rows)) if rows else \
            (HERE / f"special_t{t}_{sim}.pairs").write_text("")
        print(f"special {sim} t={t}: {len(rows)} pairs")


if __name__ == "__main__":
    main()
