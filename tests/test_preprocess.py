"""Dataset preparation (raw records -> the join's input format) and the
binary collection cache.  Pinned by hand-computed cases, the format's
invariants, and by similarity preservation: the oracle's join of the
prepared sets equals a brute-force join of the raw records computed here on
multisets (a record's repeated word counts as often as both records have
it)."""
import random
from collections import Counter

import numpy as np
import pytest

import hyset_amd as H
import oracle as O


def test_hand_example():
    c, rec = H.preprocess_text("a b a\nb c\n\nc c c b\n")
    # keys in first-appearance order: a#1 b#1 a#2 c#1 c#2 c#3; records holding
    # them 1 3 1 2 1 1 -> ids a#1=1 a#2=2 c#2=3 c#3=4 c#1=5 b#1=6
    assert c.universe == 7
    assert c.sizes.tolist() == [2, 3, 4]
    assert c.tokens.tolist() == [5, 6, 1, 2, 6, 3, 4, 5, 6]
    assert rec.tolist() == [1, 0, 3]


def test_qgrams_and_case():
    c, rec = H.preprocess_text("abcd\nABCE\n", mode="qgrams", q=2, lowercase=True)
    # padded strings give len + q - 1 grams; "abcd"/"abce" share #a, ab, bc
    assert c.sizes.tolist() == [5, 5]
    a, b = (set(c.set(i).tolist()) for i in range(2))
    assert len(a & b) == 3
    c3, _ = H.preprocess_text("aaaa", mode="qgrams", q=3)
    assert c3.sizes.tolist() == [6]  # the repeated gram "aaa" counts twice: 2 marks + 4 chars - 3 + 1 = 4 + 2


def test_integers_mode():
    c, rec = H.preprocess_text("3 1 3\n007 7\n", mode="integers")
    # "007" and "7" are one id; the second 3 of record 0 is a token of its own
    assert c.sizes.tolist() == [2, 3] and rec.tolist() == [1, 0]
    with pytest.raises(H.HysetError):
        H.preprocess_text("1 x 2\n", mode="integers")


def test_empty_input():
    c, rec = H.preprocess_text("\n\n   \n")
    assert c.n == 0 and rec.shape[0] == 0


def _corpus(n, seed, vocab=300, dup=0.3):
    rnd = random.Random(seed)
    words = [f"w{k}" for k in range(vocab)]
    weights = [1.0 / (k + 1) for k in range(vocab)]
    lines = []
    for i in range(n):
        if lines and rnd.random() < dup:
            base = rnd.choice(lines).split()
            if base and rnd.random() < 0.5:
                base[rnd.randrange(len(base))] = rnd.choice(words)
            lines.append(" ".join(base))
        else:
            lines.append(" ".join(rnd.choices(words, weights, k=rnd.randint(0, 12))))
    return lines


@pytest.mark.parametrize("seed", [1, 2])
def test_invariants(seed):
    lines = _corpus(2000, seed)
    c, rec = H.preprocess_text("\n".join(lines))
    nonempty = [i for i, l in enumerate(lines) if l.split()]
    assert sorted(rec.tolist()) == nonempty
    assert np.all(np.diff(c.sizes.astype(np.int64)) >= 0)
    freq = np.zeros(c.universe, dtype=np.int64)
    for i in range(c.n):
        s = c.set(i)
        assert np.all(np.diff(s.astype(np.int64)) > 0) and s[0] >= 1
        freq[s] += 1
        assert c.sizes[i] == len(lines[rec[i]].split())
    assert freq[0] == 0 and np.all(freq[1:] > 0)
    assert np.all(np.diff(freq[1:]) >= 0)  # ids ranked by ascending frequency


def test_join_of_prepared_sets_equals_raw_multiset_join():
    lines = _corpus(1500, 7)
    c, rec = H.preprocess_text("\n".join(lines))
    t = 0.6
    got = O.self_join(c, t)
    got_lines = {(min(rec[r], rec[s]), max(rec[r], rec[s])) for r, s in zip(got.r, got.s)}
    bags = {i: Counter(l.split()) for i, l in enumerate(lines) if l.split()}
    want = set()
    ids = sorted(bags)
    for x in range(len(ids)):
        a = bags[ids[x]]
        la = sum(a.values())
        for y in range(x):
            b = bags[ids[y]]
            lb = sum(b.values())
            ov = sum((a & b).values())
            # a self-join result: the smaller set inside the larger one's length
            # window (sets are size-sorted, so the larger one is the probe) and
            # the multiset overlap reaching minoverlap
            small, big = (la, lb) if la <= lb else (lb, la)
            b_lo, b_hi, b_mo = O.bounds("jaccard", t, big, small)[:3]
            if b_lo <= small <= b_hi and ov >= b_mo:
                want.add((ids[y], ids[x]))
    assert len(want) > 20
    assert got_lines == want


def test_binary_roundtrip_and_corruption(tmp_path):
    c = H.generate_self(H.gen_params(3000, size_dist=H.SIZE_UNIFORM, lo=1, hi=30, universe=900, zipf=1.0, seed=4))
    p = tmp_path / "c.bin"
    H.write_collection_binary(p, c)
    d = H.read_collection_binary(p)
    assert d.n == c.n and d.universe == c.universe
    assert np.array_equal(d.tokens, c.tokens) and np.array_equal(d.sizes, c.sizes)
    assert np.array_equal(d.starts, c.starts)
    raw = p.read_bytes()
    for name, bad in (("magic", b"X" + raw[1:]), ("truncated", raw[:-4]), ("trailing", raw + b"\0"),
                      ("sizes", raw[:24] + (int.from_bytes(raw[24:28], "little") + 1).to_bytes(4, "little")
                       + raw[28:])):
        q = tmp_path / f"{name}.bin"
        q.write_bytes(bad)
        with pytest.raises(H.HysetError):
            H.read_collection_binary(q)
    # a token at or above the universe
    arr = bytearray(raw)
    arr[-4:] = (c.universe + 5).to_bytes(4, "little")
    q = tmp_path / "token.bin"
    q.write_bytes(bytes(arr))
    with pytest.raises(H.HysetError):
        H.read_collection_binary(q)


def test_file_entry_points(tmp_path):
    lines = _corpus(300, 3)
    f = tmp_path / "raw.txt"
    f.write_text("\n".join(lines) + "\n")
    a, ra = H.preprocess_file(f)
    b, rb = H.preprocess_text("\n".join(lines))
    assert np.array_equal(a.tokens, b.tokens) and np.array_equal(ra, rb)
    t = tmp_path / "prepared.txt"
    H.write_collection_text(t, a)
    e = H.read_collection_text(t)
    assert np.array_equal(e.tokens, a.tokens) and np.array_equal(e.sizes, a.sizes)
