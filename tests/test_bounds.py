"""Bounds: the oracle's and the library's integer tables, pinned by the
reference's double expressions, exact rational arithmetic and the known
mismatch sums of the double form."""
from fractions import Fraction
import math

import pytest

import hyset_amd as H
import oracle as O

THRESHOLDS = [0.5, 0.6, 0.7, 0.75, 0.8, 0.85, 0.9, 0.95]


def exact_mo_jaccard(t, l1, l2):
    f = Fraction(str(t))
    return min(l1, l2, math.ceil(f * (l1 + l2) / (1 + f)))


def ref_mo_jaccard(t, l1, l2):
    # include/hyset/similarity.hpp:40-42 in Python floats (IEEE double)
    return min(l1, l2, int(math.ceil((l1 + l2) * t / (1 + t))))


def mismatch_sums(t, limit):
    return [k for k in range(2, limit + 1)
            if int(math.ceil(k * t / (1 + t))) != math.ceil(Fraction(str(t)) * k / (1 + Fraction(str(t))))]


def test_known_mismatch_sums():
    """The double quotient rounds above the exact one at these l1+l2 sums."""
    assert mismatch_sums(0.8, 400)[:7] == [63, 126, 189, 207, 252, 333, 378]
    assert mismatch_sums(0.9, 160)[:7] == [19, 38, 57, 76, 114, 133, 152]
    for t in (0.5, 0.6, 0.7, 0.75, 0.85, 0.95):
        assert mismatch_sums(t, 1400) == []


@pytest.mark.parametrize("t", THRESHOLDS)
def test_reference_mode_matches_double_formula(t):
    for l1 in range(1, 120):
        for l2 in range(1, l1 + 1):
            want = ref_mo_jaccard(t, l1, l2)
            assert O.bounds("jaccard", t, l1, l2)[2] == want
            assert H.bounds("jaccard", t, l1, l2)[2] == want


@pytest.mark.parametrize("t", [0.8, 0.9])
def test_reference_differs_from_exact_exactly_at_mismatch_sums(t):
    sums = set(mismatch_sums(t, 800))
    for l1 in range(1, 400, 3):
        for l2 in range(max(1, int(l1 * t) - 1), l1 + 1):
            ref = O.bounds("jaccard", t, l1, l2)[2]
            ex = O.bounds("jaccard", t, l1, l2, mode=O.BOUNDS_EXACT)[2]
            assert ex == exact_mo_jaccard(t, l1, l2)
            if (l1 + l2) in sums and ref < min(l1, l2):
                assert ref == ex + 1
            elif (l1 + l2) not in sums:
                assert ref == ex


def test_example_19_tokens_at_0_9():
    """Two 19-token sets sharing 18 tokens: J = 0.9 exactly; the reference's
    double bound demands 19, the exact bound 18."""
    assert O.bounds("jaccard", 0.9, 19, 19)[2] == 19
    assert O.bounds("jaccard", 0.9, 19, 19, mode=O.BOUNDS_EXACT)[2] == 18


@pytest.mark.parametrize("t", THRESHOLDS)
def test_sizes_and_prefixes_closed_form(t):
    f = Fraction(str(t))
    for l in range(1, 1401):
        mins, maxs, mo, maxp, midp = O.bounds("jaccard", t, l)
        assert mins == math.ceil(f * l)
        assert maxs == math.floor(l / f)
        assert maxp == min(l, l - mins + 1)
        assert midp == min(l, l - mo + 1)


@pytest.mark.parametrize("sim", ["cosine", "dice"])
@pytest.mark.parametrize("t", [0.6, 0.8, 0.9])
def test_cosine_dice_exact(sim, t):
    f = Fraction(str(t))
    for l1 in range(1, 80):
        for l2 in range(1, 80):
            mins, maxs, mo, _, _ = O.bounds(sim, t, l1, l2)
            if sim == "cosine":
                # smallest o with o / sqrt(l1 l2) >= t
                o = 0
                while Fraction(o * o) < f * f * l1 * l2:
                    o += 1
                assert mins == math.ceil(f * f * l1) and maxs == math.floor(l1 / (f * f))
            else:
                o = math.ceil(f * (l1 + l2) / 2)
                assert mins == math.ceil(f * l1 / (2 - f)) and maxs == math.floor((2 - f) * l1 / f)
            assert mo == min(l1, l2, o)


@pytest.mark.parametrize("sim", ["jaccard", "cosine", "dice"])
@pytest.mark.parametrize("mode", [O.BOUNDS_REFERENCE, O.BOUNDS_EXACT])
def test_library_tables_equal_oracle(sim, mode):
    for t in (0.6, 0.75, 0.8, 0.9):
        for l1 in range(1, 260, 7):
            for l2 in range(1, 260, 11):
                assert H.bounds(sim, t, l1, l2, mode) == O.bounds(sim, t, l1, l2, mode), (sim, mode, t, l1, l2)


def test_threshold_rational():
    assert O.threshold_rational(0.8) == (4, 5)
    assert O.threshold_rational(0.75) == (3, 4)
    assert O.threshold_rational(1.0) == (1, 1)
    assert O.threshold_rational(0.123) == (123, 1000)


def test_invalid_threshold_rejected():
    for bad in (0.0, -0.5, 1.5):
        with pytest.raises(H.HysetError):
            H.bounds("jaccard", bad, 10)
