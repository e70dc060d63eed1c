"""CPU oracle for the set-similarity join -- TEST INFRASTRUCTURE ONLY.

ctypes view of oracle/liboracle.so (plain-C restatement of the reference's
cpu_single / cpu_multi joins, see hs_oracle.h).  Imported by tests/, the
smoke check and the CPU-baseline leg of bench.py as the checker; the product
library never uses it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "liboracle.so"

JACCARD, COSINE, DICE = 0, 1, 2
BOUNDS_REFERENCE, BOUNDS_EXACT = 0, 1
ALLPAIRS, PPJOIN = 0, 1
SIMS = {"jaccard": JACCARD, "cosine": COSINE, "dice": DICE}


def build() -> Path:
    """Compiles the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return LIB_PATH


if not LIB_PATH.exists():
    build()

_lib = C.CDLL(str(LIB_PATH))
_u32p = C.POINTER(C.c_uint32)


class Pairs(C.Structure):
    _fields_ = [("count", C.c_uint64), ("r", _u32p), ("s", _u32p), ("overlap", _u32p),
                ("candidates", C.c_uint64), ("posting_visits", C.c_uint64), ("probe_tokens", C.c_uint64),
                ("index_entries", C.c_uint64), ("index_ns", C.c_uint64), ("join_ns", C.c_uint64)]


def _sig(name, restype, *argtypes):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)


_P = C.POINTER(Pairs)
_u32 = C.c_uint32
_sig("or_bounds", C.c_int, C.c_int, C.c_int, C.c_double, _u32, _u32, _u32p, _u32p, _u32p, _u32p, _u32p)
_sig("or_threshold_rational", C.c_int, C.c_double, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
_sig("or_self_join", C.c_int, _u32p, _u32p, _u32p, _u32, _u32, C.c_int, C.c_int, C.c_double, C.c_int, _u32,
     _u32, _P)
_sig("or_self_join_blocks", C.c_int, _u32p, _u32p, _u32p, _u32, _u32, C.c_int, C.c_int, C.c_double, C.c_int,
     _u32, _u32, _u32, _u32, _P)
_sig("or_rs_join", C.c_int, _u32p, _u32p, _u32p, _u32, _u32p, _u32p, _u32p, _u32, _u32, C.c_int, C.c_int,
     C.c_double, _P)
_sig("or_bruteforce_self", C.c_int, _u32p, _u32p, _u32p, _u32, C.c_int, C.c_int, C.c_double, _P)
_sig("or_bruteforce_rs", C.c_int, _u32p, _u32p, _u32p, _u32, _u32p, _u32p, _u32p, _u32, C.c_int, C.c_int,
     C.c_double, _P)
_sig("or_cpu_multi", C.c_int, _u32p, _u32p, _u32p, _u32, _u32, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
     _P)
_sig("or_pairs_sort", None, _P)
_sig("or_pairs_free", None, _P)
_sig("or_merge_path", _u32, _u32p, _u32, _u32p, _u32, _u32)
_sig("or_intersect", _u32, _u32p, _u32, _u32p, _u32)


def _arr(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_u32p)


def _kind(sim):
    return SIMS[sim] if isinstance(sim, str) else int(sim)


class Result:
    """Sorted (r, s, overlap) triples plus the filter counters of the run."""

    def __init__(self, p: Pairs):
        n = int(p.count)
        if n:
            self.r = np.ctypeslib.as_array(p.r, shape=(n,)).copy()
            self.s = np.ctypeslib.as_array(p.s, shape=(n,)).copy()
            self.overlap = np.ctypeslib.as_array(p.overlap, shape=(n,)).copy()
        else:
            self.r = self.s = self.overlap = np.zeros(0, dtype=np.uint32)
        self.candidates = int(p.candidates)
        self.posting_visits = int(p.posting_visits)
        self.probe_tokens = int(p.probe_tokens)
        self.index_entries = int(p.index_entries)
        # steady-clock intervals of the run (cpu_single: index build, then the join proper)
        self.index_ms = p.index_ns / 1e6
        self.join_ms = p.join_ns / 1e6

    @property
    def count(self) -> int:
        return int(self.r.shape[0])


def _run(fn, *args):
    p = Pairs()
    rc = fn(*args, C.byref(p))
    try:
        if rc != 0:
            raise RuntimeError(f"oracle call {fn.__name__} failed ({rc})")
        _lib.or_pairs_sort(C.byref(p))
        return Result(p)
    finally:
        _lib.or_pairs_free(C.byref(p))


def bounds(sim, t, l1, l2=None, mode=BOUNDS_REFERENCE):
    """(minsize, maxsize, minoverlap(l1,l2), maxprefix, midprefix) of l1."""
    l2 = l1 if l2 is None else l2
    out = [C.c_uint32() for _ in range(5)]
    rc = _lib.or_bounds(_kind(sim), mode, t, l1, l2, *[C.byref(x) for x in out])
    if rc != 0:
        raise ValueError("bad bounds arguments")
    return tuple(int(x.value) for x in out)


def threshold_rational(t):
    p, q = C.c_uint64(), C.c_uint64()
    if _lib.or_threshold_rational(t, C.byref(p), C.byref(q)) != 0:
        raise ValueError(t)
    return int(p.value), int(q.value)


def self_join(c, t, sim="jaccard", mode=BOUNDS_REFERENCE, algo=ALLPAIRS, first=0, last=None) -> Result:
    tok, tp = _arr(c.tokens)
    st, sp = _arr(c.starts)
    sz, zp = _arr(c.sizes)
    last = c.n - 1 if last is None else last
    return _run(_lib.or_self_join, tp, sp, zp, c.n, c.universe, _kind(sim), mode, t, algo, first, last)


def self_join_blocks(c, t, indexed, probe, sim="jaccard", mode=BOUNDS_REFERENCE, algo=ALLPAIRS) -> Result:
    tok, tp = _arr(c.tokens)
    st, sp = _arr(c.starts)
    sz, zp = _arr(c.sizes)
    return _run(_lib.or_self_join_blocks, tp, sp, zp, c.n, c.universe, _kind(sim), mode, t, algo, indexed[0],
                indexed[1], probe[0], probe[1])


def rs_join(R, S, t, sim="jaccard", mode=BOUNDS_REFERENCE) -> Result:
    a = [_arr(x) for x in (R.tokens, R.starts, R.sizes, S.tokens, S.starts, S.sizes)]
    uni = max(R.universe, S.universe)
    return _run(_lib.or_rs_join, a[0][1], a[1][1], a[2][1], R.n, a[3][1], a[4][1], a[5][1], S.n, uni,
                _kind(sim), mode, t)


def bruteforce_self(c, t, sim="jaccard", mode=BOUNDS_REFERENCE) -> Result:
    tok, tp = _arr(c.tokens)
    st, sp = _arr(c.starts)
    sz, zp = _arr(c.sizes)
    return _run(_lib.or_bruteforce_self, tp, sp, zp, c.n, _kind(sim), mode, t)


def bruteforce_rs(R, S, t, sim="jaccard", mode=BOUNDS_REFERENCE) -> Result:
    a = [_arr(x) for x in (R.tokens, R.starts, R.sizes, S.tokens, S.starts, S.sizes)]
    return _run(_lib.or_bruteforce_rs, a[0][1], a[1][1], a[2][1], R.n, a[3][1], a[4][1], a[5][1], S.n,
                _kind(sim), mode, t)


def cpu_multi(c, t, threads, sim="jaccard", mode=BOUNDS_REFERENCE, algo=ALLPAIRS) -> Result:
    tok, tp = _arr(c.tokens)
    st, sp = _arr(c.starts)
    sz, zp = _arr(c.sizes)
    return _run(_lib.or_cpu_multi, tp, sp, zp, c.n, c.universe, _kind(sim), mode, t, algo, threads)


def merge_path(a, b, diag) -> int:
    a, ap = _arr(a)
    b, bp = _arr(b)
    return int(_lib.or_merge_path(ap, a.shape[0], bp, b.shape[0], diag))


def intersect(a, b) -> int:
    a, ap = _arr(a)
    b, bp = _arr(b)
    return int(_lib.or_intersect(ap, a.shape[0], bp, b.shape[0]))
