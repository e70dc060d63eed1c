"""hyset_amd: exact threshold set-similarity join on MI355X (gfx950).

Python view of libhyset_amd.so (C ABI in include/hyset_amd.h).  Every join
runs in the HIP library; this module only moves arrays across the boundary.
Importing it raises if the library has not been built -- there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libhyset_amd.so"

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python -m hyset_amd.build` "
        "(hipcc for gfx950); hyset_amd has no non-HIP fallback")

_lib = C.CDLL(str(LIB_PATH))

HS_OK, HS_E_INVALID, HS_E_OOM, HS_E_OVERFLOW, HS_E_DEVICE, HS_E_UNSUPPORTED, HS_E_BUSY = 0, -1, -2, -3, -4, -5, -6
JACCARD, COSINE, DICE = 0, 1, 2
BOUNDS_REFERENCE, BOUNDS_EXACT = 0, 1
FGSS, BITMAP, HYBRID = 0, 1, 2
SIZE_UNIFORM, SIZE_LOGNORMAL = 0, 1
ALGOS = {"fgss": FGSS, "bitmap": BITMAP, "hybrid": HYBRID}
HOST_ALLPAIRS, HOST_PPJOIN = 0, 1
HOST_ALGOS = {"allpairs": HOST_ALLPAIRS, "ppjoin": HOST_PPJOIN}
SIMS = {"jaccard": JACCARD, "cosine": COSINE, "dice": DICE}

_u32p = C.POINTER(C.c_uint32)


class HostCollection(C.Structure):
    _fields_ = [("n", C.c_uint32), ("n_tokens", C.c_uint64), ("universe", C.c_uint32),
                ("tokens", _u32p), ("starts", _u32p), ("sizes", _u32p)]


class GenParams(C.Structure):
    _fields_ = [("n", C.c_uint32), ("size_dist", C.c_int32), ("size_median", C.c_double),
                ("size_sigma", C.c_double), ("size_lo", C.c_uint32), ("size_hi", C.c_uint32),
                ("universe", C.c_uint32), ("zipf_s", C.c_double), ("seed", C.c_uint64)]


class Range(C.Structure):
    _fields_ = [("first", C.c_uint32), ("last", C.c_uint32)]


class Tune(C.Structure):
    _fields_ = [("bitmap_bits", C.c_uint32), ("probes_per_tile", C.c_uint32),
                ("verify_scenario", C.c_uint32), ("host_threads", C.c_uint32),
                ("batch_candidates", C.c_uint64), ("flags", C.c_uint32), ("tile_tokens", C.c_uint32),
                ("workgroups_per_cu", C.c_uint32), ("signature_bits", C.c_uint32)]


TUNE_GRAPH = 1
TUNE_TILE_QUEUE = 2
TUNE_NO_SIGNATURE = 4
TUNE_BALANCED_TILES = 8
TUNE_STATIC_TILES = 16
TUNE_STAGE_TIMES = 32
TUNE_ITEM_MAJOR = 64
TUNE_NARROW = 128
TUNE_INDEX_MAXPREFIX = 256
PLAN_FIXED, PLAN_UPLOAD, PLAN_QUEUE, PLAN_BALANCED, PLAN_TOKEN = 1, 2, 3, 4, 5


class Result(C.Structure):
    _fields_ = [("count", C.c_uint64), ("capacity", C.c_uint64), ("r", _u32p), ("s", _u32p),
                ("overlap", _u32p), ("count_only", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("index_ms", C.c_double), ("filter_ms", C.c_double), ("transfer_ms", C.c_double),
                ("total_ms", C.c_double), ("host_filter_ms", C.c_double), ("kernel_ms", C.c_double),
                ("index_entries", C.c_uint64), ("posting_visits", C.c_uint64),
                ("candidates", C.c_uint64), ("pairs", C.c_uint64), ("probe_prefix_tokens", C.c_uint64),
                ("candidate_tokens", C.c_uint64), ("signature_tests", C.c_uint64), ("signature_bits", C.c_uint32),
                ("tile_plan", C.c_uint32), ("graph", C.c_uint32), ("reserved", C.c_uint32),
                ("candidate_probe_tokens", C.c_uint64)]


class ServerConfig(C.Structure):
    _fields_ = [("min_threshold", C.c_double), ("slots", C.c_uint32), ("hbm_budget_bytes", C.c_uint64),
                ("max_batch_queries", C.c_uint32), ("max_query_size", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("submitted", "batches", "deferred", "rejected", "max_in_flight",
                                          "max_bytes_in_flight", "signature_bits")]


def _sig(name, restype, *argtypes):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_vp = C.c_void_p
_sig("hs_generate_self", C.c_int, C.POINTER(GenParams), C.POINTER(HostCollection))
_sig("hs_generate_rs", C.c_int, C.POINTER(GenParams), C.POINTER(GenParams), C.c_double, C.c_double,
     C.c_double, C.POINTER(HostCollection), C.POINTER(HostCollection))
_sig("hs_generate_queries", C.c_int, C.POINTER(GenParams), C.POINTER(GenParams), C.POINTER(HostCollection),
     C.POINTER(HostCollection))
_sig("hs_host_collection_free", None, C.POINTER(HostCollection))
_sig("hs_read_collection_text", C.c_int, C.c_char_p, C.POINTER(HostCollection))
_sig("hs_write_collection_text", C.c_int, C.c_char_p, _u32p, _u32p, _u32p, C.c_uint32)
_sig("hs_write_pairs_text", C.c_int, C.c_char_p, _u32p, _u32p, _u32p, C.c_uint64)


class PrepParams(C.Structure):
    _fields_ = [("mode", C.c_int32), ("q", C.c_uint32), ("lowercase", C.c_int32)]


TOKENS_WORDS, TOKENS_QGRAMS, TOKENS_INTEGERS = 0, 1, 2
_TOKEN_MODES = {"words": TOKENS_WORDS, "qgrams": TOKENS_QGRAMS, "integers": TOKENS_INTEGERS}
_sig("hs_preprocess_text", C.c_int, C.c_char_p, C.c_uint64, C.POINTER(PrepParams), C.POINTER(HostCollection),
     C.POINTER(_u32p))
_sig("hs_preprocess_file", C.c_int, C.c_char_p, C.POINTER(PrepParams), C.POINTER(HostCollection), C.POINTER(_u32p))
_sig("hs_free_array", None, C.c_void_p)
_sig("hs_write_collection_binary", C.c_int, C.c_char_p, _u32p, _u32p, _u32p, C.c_uint32)
_sig("hs_read_collection_binary", C.c_int, C.c_char_p, C.POINTER(HostCollection))
_sig("hs_create", C.c_int, C.c_int, C.POINTER(_vp))
_sig("hs_destroy", None, _vp)
_sig("hs_last_error", C.c_char_p, _vp)
_sig("hs_upload_collection", C.c_int, _vp, _u32p, _u32p, _u32p, C.c_uint32, C.c_uint64, C.POINTER(_vp))
_sig("hs_free_collection", None, _vp)
_sig("hs_set_threshold", C.c_int, _vp, C.c_int32, C.c_double, C.c_int32)
_sig("hs_set_host_algorithm", C.c_int, _vp, C.c_int32)
_sig("hs_join", C.c_int, _vp, _vp, _vp, C.POINTER(Range), C.POINTER(Range), C.c_int32, C.POINTER(Tune),
     C.POINTER(Result))
_sig("hs_verify_candidates", C.c_int, _vp, _vp, _vp, _u32p, _u32p, _u32p, C.c_uint32, C.c_int32,
     C.POINTER(Result))
_sig("hs_copy_last_result", C.c_int, _vp, C.POINTER(Result))


class Coprocess(C.Structure):
    _fields_ = [("mode", C.c_int32), ("algo", C.c_int32), ("cpu_threads", C.c_uint32), ("cpu_block", C.c_uint32),
                ("device_min_claim", C.c_uint32), ("cpu_share", C.c_double), ("tune", Tune),
                ("min_device_ms", C.c_double), ("policy", C.c_int32), ("reserved", C.c_uint32)]


class CoprocessStats(C.Structure):
    _fields_ = [("device_pairs", C.c_uint64), ("host_pairs", C.c_uint64), ("device_probes", C.c_uint32),
                ("host_probes", C.c_uint32), ("device_claims", C.c_uint32), ("host_claims", C.c_uint32),
                ("split", C.c_uint32), ("host_threads", C.c_uint32), ("index_ms", C.c_double),
                ("copy_ms", C.c_double), ("device_ms", C.c_double), ("host_ms", C.c_double),
                ("wall_ms", C.c_double), ("device_alone_ms", C.c_double), ("shared_ms", C.c_double),
                ("gated", C.c_uint32), ("reserved", C.c_uint32)]


COPROC_QUEUE, COPROC_STATIC = 0, 1
_COPROC_MODES = {"queue": COPROC_QUEUE, "static": COPROC_STATIC}
POLICY_CPU_GPU, POLICY_GPU_CPU = 0, 1
_POLICIES = {"cpu,gpu": POLICY_CPU_GPU, "gpu,cpu": POLICY_GPU_CPU}
_sig("hs_join_coprocess", C.c_int, _vp, _vp, _vp, C.POINTER(Coprocess), C.POINTER(Result),
     C.POINTER(CoprocessStats))
PairSink = C.CFUNCTYPE(C.c_int, C.c_void_p, _u32p, _u32p, _u32p, C.c_uint64)
_sig("hs_stream_last_result", C.c_int, _vp, C.c_uint64, C.c_int32, PairSink, C.c_void_p)
_sig("hs_write_last_result_text", C.c_int, _vp, C.c_char_p, C.c_int32)
_sig("hs_join_host", C.c_int, _vp, _vp, _vp, C.POINTER(Range), C.POINTER(Range), C.c_uint32, C.POINTER(Result))
_u32pp = C.POINTER(_u32p)
_sig("hs_upload_batch", C.c_int, _vp, C.c_uint32, _u32pp, _u32pp, _u32pp, _u32p, C.POINTER(C.c_uint64),
     C.POINTER(_vp))
_sig("hs_join_batch", C.c_int, _vp, _vp, C.POINTER(Tune), C.POINTER(Result))
_sig("hs_autotune_batch", C.c_int, _vp, _vp, C.c_int32, C.POINTER(Tune), C.POINTER(C.c_double))
_sig("hs_get_stats", C.c_int, _vp, C.POINTER(Stats))
_sig("hs_join_noalloc", C.c_int, _vp, _vp, _vp, C.POINTER(Range), C.POINTER(Range), C.POINTER(Tune), _vp, _vp, _vp,
     C.c_uint64, C.POINTER(C.c_uint64))
NOALLOC_SCRATCH_BYTES = 128 << 20
_sig("hs_copy_bandwidth", C.c_int, _vp, C.c_uint64, C.c_int32, C.POINTER(C.c_double))
_sig("hs_popcount_rate", C.c_int, _vp, C.c_int32, C.POINTER(C.c_double))
_sig("hs_server_create", C.c_int, C.c_int, _vp, C.c_int32, C.c_int32, C.POINTER(ServerConfig), C.POINTER(_vp))
_sig("hs_server_submit", C.c_int, _vp, _u32p, _u32p, C.c_uint32, C.c_uint64, C.c_double, C.POINTER(C.c_uint64))
_sig("hs_server_wait", C.c_int, _vp, C.c_uint64, C.POINTER(Result), C.POINTER(C.c_double))
_sig("hs_server_get_stats", C.c_int, _vp, C.POINTER(ServerStats))
_sig("hs_server_destroy", None, _vp)
_sig("hs_bounds", C.c_int, C.c_int32, C.c_double, C.c_int32, C.c_uint32, C.c_uint32, _u32p)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_sig("hs_features", C.c_int, _vp, _vp, C.POINTER(Range), _f32p)
_sig("hs_linear_fit", C.c_int, _f64p, _f64p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, _f64p)
_sig("hs_linear_predict", C.c_int, _f64p, _f64p, C.c_uint32, C.c_uint32, C.c_uint32, _f64p)
_sig("hs_mlp_create", C.c_int, _vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(_vp))
_sig("hs_mlp_destroy", None, _vp)
_sig("hs_mlp_num_params", C.c_uint32, _vp)
_sig("hs_mlp_set_normalization", C.c_int, _vp, _f32p, _f32p)
_sig("hs_mlp_get_params", C.c_int, _vp, _f32p)
_sig("hs_mlp_set_params", C.c_int, _vp, _f32p)
_sig("hs_mlp_predict", C.c_int, _vp, _f32p, C.c_uint32, _f32p)
_sig("hs_mlp_loss_grad", C.c_int, _vp, _f32p, _f32p, C.c_uint32, _f32p, _f32p)
_sig("hs_mlp_train", C.c_int, _vp, _f32p, _f32p, C.c_uint32, C.c_uint32, C.c_float, _f32p)
_sig("hs_split_run", C.c_int, _vp, C.c_uint32, C.POINTER(_vp), C.POINTER(C.c_int32), C.c_uint32,
     C.POINTER(C.c_uint64), _f64p, _f64p)

class CostCoefficients(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("intercept", "set_size_std", "average_set_size", "token_freq", "size",
                                          "size_square", "universe_size", "index_size", "bitmap")] + \
               [("enabled", C.c_int32), ("reserved", C.c_uint32)]


class PartitionPair(C.Structure):
    _fields_ = [("left", Range), ("right", Range), ("technique", C.c_int32), ("reserved", C.c_uint32),
                ("pairs", C.c_uint64), ("ms", C.c_double)]


COST_TERMS = ("intercept", "set_size_std", "average_set_size", "token_freq", "size", "size_square",
              "universe_size", "index_size", "bitmap")
COST_WRAP_SQUARE = 1
_sig("hs_features_pair", C.c_int, _vp, _vp, C.POINTER(Range), C.POINTER(Range), _f32p)
_sig("hs_partition_even", C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Range), C.POINTER(C.c_uint32))
_sig("hs_partition_dichotomize", C.c_int, C.c_uint32, C.c_uint32, C.c_double, C.POINTER(Range))
_sig("hs_cost_run", C.c_int, _vp, _vp, C.POINTER(Range), C.c_uint32, C.POINTER(CostCoefficients), C.c_uint32,
     C.c_uint32, C.POINTER(C.c_int32), _f32p, _f64p)
_sig("hs_run_partition_pairs", C.c_int, _vp, _vp, C.POINTER(PartitionPair), C.c_uint32, C.c_uint32,
     C.POINTER(Tune), C.POINTER(Result), _f64p)


def partition_even(n, block, parts):
    """partitioner::even (hs_partition_even): set-id ranges [(first, last)]."""
    out = (Range * max(1, int(parts)))()
    k = C.c_uint32()
    rc = _lib.hs_partition_even(int(n), int(block), int(parts), out, C.byref(k))
    if rc != HS_OK:
        raise HysetError(rc, "hs_partition_even: invalid arguments")
    return [(out[i].first, out[i].last) for i in range(k.value)]


def partition_dichotomize(n, block, split_pct):
    """partitioner::dichotomize (hs_partition_dichotomize): (whole, Rs, Rl), None for an empty part."""
    out = (Range * 3)()
    rc = _lib.hs_partition_dichotomize(int(n), int(block), float(split_pct), out)
    if rc != HS_OK:
        raise HysetError(rc, "hs_partition_dichotomize: invalid arguments")
    return tuple((r.first, r.last) if r.first <= r.last else None for r in out)


def cost_coefficients(coefs, enabled=(1, 1, 1, 1)):
    """A (4, 9) array (rows HS_TECH_*, columns COST_TERMS) as the C array of 4 hs_cost_coefficients."""
    coefs = np.asarray(coefs, dtype=np.float64)
    assert coefs.shape == (4, len(COST_TERMS))
    arr = (CostCoefficients * 4)()
    for k in range(4):
        for j, name in enumerate(COST_TERMS):
            setattr(arr[k], name, float(coefs[k, j]))
        arr[k].enabled = int(bool(enabled[k]))
    return arr


NFEATURES = 7
FEATURES = ("size", "averageSetSize", "setSizeStd", "tokenFreqDistribution", "universeSize", "indexSize",
            "threshold")
TECH_FGSS, TECH_BITMAP, TECH_HYBRID, TECH_CPU = 0, 1, 2, 3
TECHNIQUES = ("fgss", "bitmap", "hybrid", "cpu")


def bounds(sim, t, l1, l2=None, mode=BOUNDS_REFERENCE):
    """(minsize, maxsize, minoverlap(l1,l2), maxprefix, midprefix) of l1 as the device tables hold them."""
    l2 = l1 if l2 is None else l2
    out = np.zeros(5, dtype=np.uint32)
    kind = SIMS[sim] if isinstance(sim, str) else int(sim)
    rc = _lib.hs_bounds(kind, float(t), int(mode), l1, l2, _ptr(out))
    if rc != HS_OK:
        raise HysetError(rc, "hs_bounds")
    return tuple(int(x) for x in out)


class HysetError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hyset_amd error {code}: {msg}")
        self.code = code


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


# ----------------------------------------------------------------- inputs

@dataclass
class SetCollection:
    """Host CSR collection: set i = tokens[starts[i]:starts[i]+sizes[i]]."""
    tokens: np.ndarray
    starts: np.ndarray
    sizes: np.ndarray
    universe: int

    @property
    def n(self) -> int:
        return int(self.sizes.shape[0])

    def set(self, i: int) -> np.ndarray:
        return self.tokens[self.starts[i]:self.starts[i] + self.sizes[i]]

    @staticmethod
    def from_sets(sets) -> "SetCollection":
        sizes = np.array([len(s) for s in sets], dtype=np.uint32)
        starts = np.zeros(len(sets), dtype=np.uint32)
        if len(sets):
            starts[1:] = np.cumsum(sizes[:-1], dtype=np.uint64).astype(np.uint32)
        toks = np.array([t for s in sets for t in s], dtype=np.uint32)
        uni = int(toks.max()) + 1 if toks.size else 1
        return SetCollection(toks, starts, sizes, uni)


def _take(hc: HostCollection) -> SetCollection:
    try:
        n, nt = hc.n, hc.n_tokens
        toks = np.ctypeslib.as_array(hc.tokens, shape=(max(nt, 1),))[:nt].copy()
        starts = np.ctypeslib.as_array(hc.starts, shape=(max(n, 1),))[:n].copy()
        sizes = np.ctypeslib.as_array(hc.sizes, shape=(max(n, 1),))[:n].copy()
        return SetCollection(toks, starts, sizes, int(hc.universe))
    finally:
        _lib.hs_host_collection_free(C.byref(hc))


def gen_params(n, *, size_dist=SIZE_LOGNORMAL, median=10.0, sigma=0.6, lo=2, hi=200, universe=1_000_000,
               zipf=1.1, seed=1) -> GenParams:
    return GenParams(n, size_dist, float(median), float(sigma), lo, hi, universe, float(zipf), seed)


def generate_self(p: GenParams) -> SetCollection:
    hc = HostCollection()
    rc = _lib.hs_generate_self(C.byref(p), C.byref(hc))
    if rc != HS_OK:
        raise HysetError(rc, "hs_generate_self")
    return _take(hc)


def generate_rs(pr: GenParams, ps: GenParams, dup_frac=0.2, edit_lo=0.05, edit_hi=0.15):
    hr, hs_ = HostCollection(), HostCollection()
    rc = _lib.hs_generate_rs(C.byref(pr), C.byref(ps), dup_frac, edit_lo, edit_hi, C.byref(hr), C.byref(hs_))
    if rc != HS_OK:
        raise HysetError(rc, "hs_generate_rs")
    return _take(hr), _take(hs_)


def generate_queries(pi: GenParams, pq: GenParams):
    hi, hq = HostCollection(), HostCollection()
    rc = _lib.hs_generate_queries(C.byref(pi), C.byref(pq), C.byref(hi), C.byref(hq))
    if rc != HS_OK:
        raise HysetError(rc, "hs_generate_queries")
    return _take(hi), _take(hq)


def read_collection_text(path) -> SetCollection:
    hc = HostCollection()
    rc = _lib.hs_read_collection_text(str(path).encode(), C.byref(hc))
    if rc != HS_OK:
        raise HysetError(rc, f"cannot read {path}")
    return _take(hc)


def write_collection_text(path, c: SetCollection) -> None:
    t, st, sz = _u32(c.tokens), _u32(c.starts), _u32(c.sizes)
    rc = _lib.hs_write_collection_text(str(path).encode(), _ptr(t), _ptr(st), _ptr(sz), c.n)
    if rc != HS_OK:
        raise HysetError(rc, f"cannot write {path}")


def _prep(mode, q, lowercase):
    return PrepParams(_TOKEN_MODES[mode] if isinstance(mode, str) else int(mode), int(q), int(bool(lowercase)))


def _take_records(hc: HostCollection, rec) -> tuple:
    try:
        n = hc.n
        lines = np.ctypeslib.as_array(rec, shape=(max(n, 1),))[:n].copy() if rec else np.zeros(0, np.uint32)
    finally:
        if rec:
            _lib.hs_free_array(rec)
    return _take(hc), lines


def preprocess_text(text, mode="words", q=3, lowercase=False):
    """Raw records (one per line) -> (collection, record_of): tokens by
    word, character q-gram or integer id, repeats numbered, ids ranked by
    ascending frequency from 1, sets sorted by size; record_of[i] is the line
    of set i (hs_preprocess_text)."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    hc, rec = HostCollection(), _u32p()
    p = _prep(mode, q, lowercase)
    rc = _lib.hs_preprocess_text(data, len(data), C.byref(p), C.byref(hc), C.byref(rec))
    if rc != HS_OK:
        raise HysetError(rc, "hs_preprocess_text")
    return _take_records(hc, rec)


def preprocess_file(path, mode="words", q=3, lowercase=False):
    hc, rec = HostCollection(), _u32p()
    p = _prep(mode, q, lowercase)
    rc = _lib.hs_preprocess_file(str(path).encode(), C.byref(p), C.byref(hc), C.byref(rec))
    if rc != HS_OK:
        raise HysetError(rc, f"cannot preprocess {path}")
    return _take_records(hc, rec)


def write_collection_binary(path, c: SetCollection) -> None:
    t, st, sz = _u32(c.tokens), _u32(c.starts), _u32(c.sizes)
    rc = _lib.hs_write_collection_binary(str(path).encode(), _ptr(t), _ptr(st), _ptr(sz), c.n)
    if rc != HS_OK:
        raise HysetError(rc, f"cannot write {path}")


def read_collection_binary(path) -> SetCollection:
    hc = HostCollection()
    rc = _lib.hs_read_collection_binary(str(path).encode(), C.byref(hc))
    if rc != HS_OK:
        raise HysetError(rc, f"cannot read {path}")
    return _take(hc)


def write_pairs_text(path, r, s, overlap=None) -> None:
    r, s = _u32(r), _u32(s)
    ov = _u32(overlap) if overlap is not None else None
    rc = _lib.hs_write_pairs_text(str(path).encode(), _ptr(r), _ptr(s), _ptr(ov) if ov is not None else None,
                                  r.shape[0])
    if rc != HS_OK:
        raise HysetError(rc, f"cannot write {path}")


# ------------------------------------------------------------------ device

class Context:
    """One HIP stream plus its scratch on one device (hs_ctx)."""

    def __init__(self, device: int = 0):
        h = _vp()
        rc = _lib.hs_create(device, C.byref(h))
        if rc != HS_OK:
            raise HysetError(rc, f"hs_create({device}) failed")
        self._h = h
        self._colls = []
        self._last_pairs = 0

    def _check(self, rc: int, what: str):
        if rc != HS_OK:
            raise HysetError(rc, f"{what}: {_lib.hs_last_error(self._h).decode(errors='replace')}")

    def upload(self, c: SetCollection) -> "DeviceCollection":
        t, st, sz = _u32(c.tokens), _u32(c.starts), _u32(c.sizes)
        h = _vp()
        self._check(_lib.hs_upload_collection(self._h, _ptr(t), _ptr(st), _ptr(sz), c.n, t.shape[0], C.byref(h)),
                    "hs_upload_collection")
        dc = DeviceCollection(h, c)
        self._colls.append(dc)
        return dc

    def set_threshold(self, t: float, sim="jaccard", bounds=BOUNDS_REFERENCE) -> None:
        kind = SIMS[sim] if isinstance(sim, str) else int(sim)
        self._check(_lib.hs_set_threshold(self._h, kind, float(t), int(bounds)), "hs_set_threshold")

    def set_host_algorithm(self, algorithm="allpairs") -> None:
        """The host cores' filtering (hs_set_host_algorithm): "allpairs" or
        "ppjoin", for the hybrid technique's filter, join_host and the host
        side of join_coprocess."""
        a = HOST_ALGOS[algorithm] if isinstance(algorithm, str) else int(algorithm)
        self._check(_lib.hs_set_host_algorithm(self._h, a), "hs_set_host_algorithm")

    def join(self, indexed: "DeviceCollection", probe: "DeviceCollection" = None, *, algo="fgss",
             indexed_range=None, probe_range=None, tune: Tune | None = None, count_only=False,
             capacity=None):
        """Returns (r, s, overlap) arrays, or the pair count when count_only.
        The arrays are passed to hs_join sized for `capacity` pairs (default:
        a quarter more than the last join returned, at least 4096), so a
        result that fits reaches the host in the join's own round trip; a
        larger one is fetched afterwards (hs_copy_last_result)."""
        probe = indexed if probe is None else probe
        a = ALGOS[algo] if isinstance(algo, str) else int(algo)
        ir = Range(*indexed_range) if indexed_range is not None else None
        pr = Range(*probe_range) if probe_range is not None else None
        args = (self._h, indexed._h, probe._h, C.byref(ir) if ir else None, C.byref(pr) if pr else None, a,
                C.byref(tune) if tune else None)
        if count_only:
            res = Result()
            res.count_only = 1
            self._check(_lib.hs_join(*args, C.byref(res)), "hs_join")
            return int(res.count)
        cap = int(capacity) if capacity is not None else max(4096, self._last_pairs + self._last_pairs // 4)
        r = np.empty(max(cap, 1), dtype=np.uint32)
        s = np.empty_like(r)
        ov = np.empty_like(r)
        res = Result(0, cap, _ptr(r), _ptr(s), _ptr(ov), 0)
        rc = _lib.hs_join(*args, C.byref(res))
        n = int(res.count)
        self._last_pairs = n
        if rc == HS_E_OVERFLOW:
            return self.last_result(n)
        self._check(rc, "hs_join")
        return r[:n], s[:n], ov[:n]

    def join_noalloc(self, indexed: "DeviceCollection", probe: "DeviceCollection" = None, *, r: int, s: int,
                     overlap: int, capacity: int, indexed_range=None, probe_range=None,
                     tune: Tune | None = None) -> int:
        """hs_join_noalloc: the fgss join into caller-owned DEVICE arrays (r,
        s, overlap are device addresses of `capacity` uint32 each), with no
        allocation by the library; returns the pair count.  Too many pairs
        raise HysetError(HS_E_OVERFLOW) whose .count is the number needed."""
        probe = indexed if probe is None else probe
        ir = Range(*indexed_range) if indexed_range is not None else None
        pr = Range(*probe_range) if probe_range is not None else None
        n = C.c_uint64()
        rc = _lib.hs_join_noalloc(self._h, indexed._h, probe._h, C.byref(ir) if ir else None,
                                  C.byref(pr) if pr else None, C.byref(tune) if tune else None, r, s, overlap,
                                  int(capacity), C.byref(n))
        if rc == HS_E_OVERFLOW:
            e = HysetError(rc, f"hs_join_noalloc: {n.value} pairs, capacity {capacity}")
            e.count = int(n.value)
            raise e
        self._check(rc, "hs_join_noalloc")
        return int(n.value)

    def join_coprocess(self, indexed: "DeviceCollection", probe: "DeviceCollection" = None, *, mode="queue",
                       algo="fgss", cpu_threads=0, cpu_block=0, device_min_claim=0, cpu_share=0.0,
                       tune: Tune | None = None, count_only=False, min_device_ms=0.0, policy="cpu,gpu"):
        """One join shared by the device and host threads (hs_join_coprocess).
        min_device_ms: the queue mode's break-even gate (0: the library's
        default, < 0: always share).  policy (static mode): "cpu,gpu" or
        "gpu,cpu", the sides of the small part Rs (cpu_share of the probe
        sets) and of the large part, as the reference's --policy pair; algo
        "hybrid" makes the device side the cpu-gpu technique.  Returns
        ((r, s, overlap) or the pair count, stats dict)."""
        probe = indexed if probe is None else probe
        opt = Coprocess(_COPROC_MODES[mode] if isinstance(mode, str) else int(mode),
                        ALGOS[algo] if isinstance(algo, str) else int(algo), int(cpu_threads), int(cpu_block),
                        int(device_min_claim), float(cpu_share), tune if tune is not None else Tune(),
                        float(min_device_ms), _POLICIES[policy] if isinstance(policy, str) else int(policy), 0)
        res = Result()
        res.count_only = 1
        st = CoprocessStats()
        self._check(_lib.hs_join_coprocess(self._h, indexed._h, probe._h, C.byref(opt), C.byref(res), C.byref(st)),
                    "hs_join_coprocess")
        info = {k: getattr(st, k) for k, _ in CoprocessStats._fields_}
        if count_only:
            return int(res.count), info
        return self.last_result(int(res.count)), info

    def join_host(self, indexed: "DeviceCollection", probe: "DeviceCollection" = None, *, threads=1,
                  indexed_range=None, probe_range=None):
        """The CPU partition (hs_join_host): (r, s, overlap) computed on host cores."""
        probe = indexed if probe is None else probe
        ir = Range(*indexed_range) if indexed_range is not None else None
        pr = Range(*probe_range) if probe_range is not None else None
        cap = 1 << 16
        while True:
            r = np.empty(cap, dtype=np.uint32)
            s = np.empty_like(r)
            ov = np.empty_like(r)
            res = Result(0, cap, _ptr(r), _ptr(s), _ptr(ov), 0)
            rc = _lib.hs_join_host(self._h, indexed._h, probe._h, C.byref(ir) if ir else None,
                                   C.byref(pr) if pr else None, threads, C.byref(res))
            if rc == HS_E_OVERFLOW:
                cap = int(res.count)
                continue
            self._check(rc, "hs_join_host")
            n = int(res.count)
            return r[:n], s[:n], ov[:n]

    def last_result(self, count: int):
        r = np.empty(max(count, 1), dtype=np.uint32)
        s = np.empty_like(r)
        ov = np.empty_like(r)
        res = Result(0, count, _ptr(r), _ptr(s), _ptr(ov), 0)
        self._check(_lib.hs_copy_last_result(self._h, C.byref(res)), "hs_copy_last_result")
        return r[:count], s[:count], ov[:count]

    def stream_last_result(self, consume, *, chunk=0, overlap=True):
        """Feeds the pairs of the last join to consume(r, s, overlap or None)
        chunk by chunk (hs_stream_last_result); the arrays are copies.  A
        falsy return from consume continues, anything else stops the
        stream with HysetError."""
        def sink(_user, r, s, ov, n):
            n = int(n)
            view = lambda p: np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.empty(0, np.uint32)
            try:
                stop = consume(view(r), view(s), view(ov) if ov else None)
            except Exception:  # an exception must not cross the C frame
                return 1
            return 1 if stop else 0
        cb = PairSink(sink)
        self._check(_lib.hs_stream_last_result(self._h, int(chunk), 1 if overlap else 0, cb, None),
                    "hs_stream_last_result")

    def write_last_result_text(self, path, *, overlap=False) -> None:
        """Streams the pairs of the last join into a text file of "r s" (or
        "r s overlap") lines."""
        self._check(_lib.hs_write_last_result_text(self._h, str(path).encode(), 1 if overlap else 0),
                    "hs_write_last_result_text")

    def upload_batch(self, colls) -> "DeviceCollection":
        """One batched device collection from a list of SetCollections."""
        arrs = [(_u32(c.tokens), _u32(c.starts), _u32(c.sizes)) for c in colls]
        n = len(colls)
        toks = (_u32p * n)(*[_ptr(a[0]) for a in arrs])
        sts = (_u32p * n)(*[_ptr(a[1]) for a in arrs])
        szs = (_u32p * n)(*[_ptr(a[2]) for a in arrs])
        nsets = np.array([c.n for c in colls], dtype=np.uint32)
        ntok = np.array([a[0].shape[0] for a in arrs], dtype=np.uint64)
        h = _vp()
        self._check(_lib.hs_upload_batch(self._h, n, toks, sts, szs, _ptr(nsets),
                                         ntok.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(h)), "hs_upload_batch")
        dc = DeviceCollection(h, None)
        dc.join_first = np.concatenate([[0], np.cumsum(nsets.astype(np.int64))])
        self._colls.append(dc)
        return dc

    def join_batch(self, batch: "DeviceCollection", tune: Tune | None = None, count_only=False):
        """Self-joins every collection of a batch; pairs are in batch ids."""
        res = Result()
        res.count_only = 1
        self._check(_lib.hs_join_batch(self._h, batch._h, C.byref(tune) if tune else None, C.byref(res)),
                    "hs_join_batch")
        if count_only:
            return int(res.count)
        return self.last_result(int(res.count))

    def autotune_batch(self, batch: "DeviceCollection", reps=3):
        t = Tune()
        ms = C.c_double()
        self._check(_lib.hs_autotune_batch(self._h, batch._h, reps, C.byref(t), C.byref(ms)), "hs_autotune_batch")
        return t, ms.value

    def verify_candidates(self, indexed, probe, probes, offsets, candidates, scenario=0):
        pr, of, ca = _u32(probes), _u32(offsets), _u32(candidates)
        res = Result()
        res.count_only = 1
        self._check(_lib.hs_verify_candidates(self._h, indexed._h, probe._h, _ptr(pr), _ptr(of), _ptr(ca),
                                              pr.shape[0], scenario, C.byref(res)), "hs_verify_candidates")
        return self.last_result(int(res.count))

    def features(self, coll: "DeviceCollection", rng=None) -> np.ndarray:
        """Partition features of sets [rng] (hs_features) under the current threshold."""
        out = np.zeros(NFEATURES, dtype=np.float32)
        r = Range(*rng) if rng is not None else None
        self._check(_lib.hs_features(self._h, coll._h, C.byref(r) if r else None,
                                     out.ctypes.data_as(_f32p)), "hs_features")
        return out

    def features_pair(self, coll: "DeviceCollection", left, right) -> np.ndarray:
        """statistics::extract(left, right) of two set ranges (hs_features_pair)."""
        out = np.zeros(NFEATURES, dtype=np.float32)
        self._check(_lib.hs_features_pair(self._h, coll._h, C.byref(Range(*left)), C.byref(Range(*right)),
                                          out.ctypes.data_as(_f32p)), "hs_features_pair")
        return out

    def cost_run(self, coll: "DeviceCollection", parts, coefs, enabled=(1, 1, 1, 1), bitmap_bits=64,
                 wrap_square=False):
        """cost::run over the partition pairs (i, j), i <= j (hs_cost_run).
        coefs: (4, 9) per HS_TECH_* in COST_TERMS order.  Returns (pairs
        [(i, j)], techniques, features (P, 7), costs (P, 4))."""
        n = len(parts)
        rng = (Range * n)(*[Range(*p) for p in parts])
        npairs = n * (n + 1) // 2
        tech = np.zeros(npairs, dtype=np.int32)
        feats = np.zeros((npairs, NFEATURES), dtype=np.float32)
        costs = np.zeros((npairs, 4), dtype=np.float64)
        self._check(_lib.hs_cost_run(self._h, coll._h, rng, n, cost_coefficients(coefs, enabled), int(bitmap_bits),
                                     COST_WRAP_SQUARE if wrap_square else 0,
                                     tech.ctypes.data_as(C.POINTER(C.c_int32)), feats.ctypes.data_as(_f32p),
                                     costs.ctypes.data_as(_f64p)), "hs_cost_run")
        ij = [(i, j) for i in range(n) for j in range(i, n)]
        return ij, tech, feats, costs

    def run_partition_pairs(self, coll: "DeviceCollection", plan, *, cpu_threads=1, tune: Tune | None = None,
                            count_only=False):
        """Runs partition pairs [(left, right, technique)] of coll's self-join
        (hs_run_partition_pairs).  Returns (triples or the pair count,
        per-pair pairs, per-pair ms, wall ms)."""
        n = len(plan)
        pp = (PartitionPair * max(n, 1))()
        for k, (left, right, tech) in enumerate(plan):
            pp[k].left = Range(*left)
            pp[k].right = Range(*right)
            pp[k].technique = int(tech)
        wall = C.c_double()
        res = Result()
        res.count_only = int(bool(count_only))
        rc = _lib.hs_run_partition_pairs(self._h, coll._h, pp, n, int(cpu_threads), C.byref(tune) if tune else None,
                                         C.byref(res), C.byref(wall))
        if rc != HS_E_OVERFLOW:
            self._check(rc, "hs_run_partition_pairs")
        counts = np.array([pp[k].pairs for k in range(n)], dtype=np.uint64)
        ms = np.array([pp[k].ms for k in range(n)], dtype=np.float64)
        if count_only:
            return int(res.count), counts, ms, wall.value
        return self.last_result(int(res.count)), counts, ms, wall.value

    def split_run(self, colls, techniques, cpu_threads=1):
        """Runs every collection's self-join with its technique (hs_split_run);
        returns (counts, per-unit ms, wall ms)."""
        n = len(colls)
        hs = (_vp * n)(*[c._h for c in colls])
        tech = np.ascontiguousarray(techniques, dtype=np.int32)
        assert tech.shape == (n,)
        counts = np.zeros(n, dtype=np.uint64)
        unit_ms = np.zeros(n, dtype=np.float64)
        wall = C.c_double()
        self._check(_lib.hs_split_run(self._h, n, hs, tech.ctypes.data_as(C.POINTER(C.c_int32)), cpu_threads,
                                      counts.ctypes.data_as(C.POINTER(C.c_uint64)), unit_ms.ctypes.data_as(_f64p),
                                      C.byref(wall)), "hs_split_run")
        return counts, unit_ms, wall.value

    def mlp(self, n_in, h1, h2, n_out, seed=1) -> "Mlp":
        return Mlp(self, n_in, h1, h2, n_out, seed)

    def stats(self) -> dict:
        st = Stats()
        _lib.hs_get_stats(self._h, C.byref(st))
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def popcount_rate(self, reps=10) -> float:
        """Giga-operations/s of XOR-popcount chains (hs_popcount_rate)."""
        g = C.c_double()
        self._check(_lib.hs_popcount_rate(self._h, int(reps), C.byref(g)), "hs_popcount_rate")
        return g.value

    def copy_bandwidth(self, nbytes=1 << 30, reps=10) -> float:
        """GB/s (read + write) of a plain device copy kernel."""
        g = C.c_double()
        self._check(_lib.hs_copy_bandwidth(self._h, nbytes, reps, C.byref(g)), "hs_copy_bandwidth")
        return g.value

    def close(self):
        for c in self._colls:
            c.free()
        self._colls = []
        if self._h:
            _lib.hs_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceCollection:
    def __init__(self, h, host: SetCollection):
        self._h = h
        self.host = host

    @property
    def n(self):
        return self.host.n if self.host is not None else int(self.join_first[-1])

    def free(self):
        if self._h:
            _lib.hs_free_collection(self._h)
            self._h = None


def _f32(a, cols) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 1:
        a = a.reshape(-1, cols)
    assert a.shape[1] == cols, (a.shape, cols)
    return a


class Mlp:
    """The learned cost model's network (hs_mlp): trained and evaluated by
    the library's device kernels."""

    def __init__(self, ctx: Context, n_in, h1, h2, n_out, seed=1):
        h = _vp()
        ctx._check(_lib.hs_mlp_create(ctx._h, n_in, h1, h2, n_out, seed, C.byref(h)), "hs_mlp_create")
        self._ctx, self._h = ctx, h
        self.shape = (n_in, h1, h2, n_out)

    def _check(self, rc, what):
        self._ctx._check(rc, what)

    @property
    def num_params(self) -> int:
        return int(_lib.hs_mlp_num_params(self._h))

    def set_normalization(self, mean, stdv):
        m = np.ascontiguousarray(mean, dtype=np.float32)
        s = np.ascontiguousarray(stdv, dtype=np.float32)
        assert m.shape == s.shape == (self.shape[0],)
        self._check(_lib.hs_mlp_set_normalization(self._h, m.ctypes.data_as(_f32p), s.ctypes.data_as(_f32p)),
                    "hs_mlp_set_normalization")

    def params(self) -> np.ndarray:
        out = np.empty(self.num_params, dtype=np.float32)
        self._check(_lib.hs_mlp_get_params(self._h, out.ctypes.data_as(_f32p)), "hs_mlp_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (self.num_params,)
        self._check(_lib.hs_mlp_set_params(self._h, p.ctypes.data_as(_f32p)), "hs_mlp_set_params")

    def predict(self, X) -> np.ndarray:
        X = _f32(X, self.shape[0])
        Y = np.empty((X.shape[0], self.shape[3]), dtype=np.float32)
        self._check(_lib.hs_mlp_predict(self._h, X.ctypes.data_as(_f32p), X.shape[0], Y.ctypes.data_as(_f32p)),
                    "hs_mlp_predict")
        return Y

    def loss_grad(self, X, Y):
        X, Y = _f32(X, self.shape[0]), _f32(Y, self.shape[3])
        loss = C.c_float()
        g = np.empty(self.num_params, dtype=np.float32)
        self._check(_lib.hs_mlp_loss_grad(self._h, X.ctypes.data_as(_f32p), Y.ctypes.data_as(_f32p), X.shape[0],
                                          C.byref(loss), g.ctypes.data_as(_f32p)), "hs_mlp_loss_grad")
        return loss.value, g

    def train(self, X, Y, epochs=2000, lr=1e-2) -> float:
        X, Y = _f32(X, self.shape[0]), _f32(Y, self.shape[3])
        loss = C.c_float()
        self._check(_lib.hs_mlp_train(self._h, X.ctypes.data_as(_f32p), Y.ctypes.data_as(_f32p), X.shape[0],
                                      epochs, lr, C.byref(loss)), "hs_mlp_train")
        return loss.value

    def close(self):
        if self._h:
            _lib.hs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def linear_fit(X, Y, ridge=0.0) -> np.ndarray:
    """Ridge least squares (hs_linear_fit): coef of shape (d + 1, n_out), last row = intercepts."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    if Y.ndim == 1:
        Y = Y[:, None]
    n, d = X.shape
    coef = np.empty((d + 1, Y.shape[1]), dtype=np.float64)
    rc = _lib.hs_linear_fit(X.ctypes.data_as(_f64p), Y.ctypes.data_as(_f64p), n, d, Y.shape[1], ridge,
                            coef.ctypes.data_as(_f64p))
    if rc != HS_OK:
        raise HysetError(rc, "hs_linear_fit: singular or invalid system")
    return coef


def linear_predict(coef, X) -> np.ndarray:
    """|X coef + intercept| per output (hs_linear_predict)."""
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, d = X.shape
    assert coef.shape[0] == d + 1
    Y = np.empty((n, coef.shape[1]), dtype=np.float64)
    rc = _lib.hs_linear_predict(coef.ctypes.data_as(_f64p), X.ctypes.data_as(_f64p), n, d, coef.shape[1],
                                Y.ctypes.data_as(_f64p))
    if rc != HS_OK:
        raise HysetError(rc, "hs_linear_predict")
    return Y


def sort_pairs(r, s, ov):
    """Sorts result triples by (r, s)."""
    key = (np.asarray(r, dtype=np.uint64) << np.uint64(32)) | np.asarray(s, dtype=np.uint64)
    order = np.argsort(key, kind="stable")
    return np.asarray(r)[order], np.asarray(s)[order], np.asarray(ov)[order]


class Server:
    """Continuous-batching query server over a device-resident index."""

    def __init__(self, index: DeviceCollection, device=0, sim="jaccard", bounds=BOUNDS_REFERENCE,
                 min_threshold=0.6, slots=4, hbm_budget_bytes=1 << 30, max_batch_queries=65536, max_query_size=0):
        cfg = ServerConfig(min_threshold, slots, hbm_budget_bytes, max_batch_queries, max_query_size)
        h = _vp()
        kind = SIMS[sim] if isinstance(sim, str) else int(sim)
        rc = _lib.hs_server_create(device, index._h, kind, int(bounds), C.byref(cfg), C.byref(h))
        if rc != HS_OK:
            raise HysetError(rc, "hs_server_create")
        self._h = h
        self._keep = index

    def submit(self, queries: SetCollection, threshold: float) -> int:
        t, sz = _u32(queries.tokens), _u32(queries.sizes)
        rid = C.c_uint64()
        rc = _lib.hs_server_submit(self._h, _ptr(t), _ptr(sz), queries.n, t.shape[0], float(threshold), C.byref(rid))
        if rc != HS_OK:
            raise HysetError(rc, "hs_server_submit")
        return int(rid.value)

    def wait(self, rid: int, count_only=False):
        """(r, s, overlap, latency_ms), or (count, latency_ms) when count_only."""
        res = Result()
        res.count_only = 1
        lat = C.c_double()
        rc = _lib.hs_server_wait(self._h, rid, C.byref(res), C.byref(lat))
        if rc != HS_OK:
            raise HysetError(rc, "hs_server_wait")
        return int(res.count), lat.value

    def wait_pairs(self, rid: int):
        """(r, s, overlap, latency_ms) of a request; r indexes its query sets."""
        res = Result(0, 0, None, None, None, 0)
        lat = C.c_double()
        rc = _lib.hs_server_wait(self._h, rid, C.byref(res), C.byref(lat))
        if rc == HS_E_OVERFLOW:
            n = int(res.count)
            r = np.empty(n, dtype=np.uint32)
            s = np.empty_like(r)
            ov = np.empty_like(r)
            res = Result(0, n, _ptr(r), _ptr(s), _ptr(ov), 0)
            rc = _lib.hs_server_wait(self._h, rid, C.byref(res), C.byref(lat))
            if rc != HS_OK:
                raise HysetError(rc, "hs_server_wait")
            return r, s, ov, lat.value
        if rc != HS_OK:
            raise HysetError(rc, "hs_server_wait")
        z = np.zeros(0, dtype=np.uint32)
        return z, z.copy(), z.copy(), lat.value

    def stats(self) -> dict:
        st = ServerStats()
        _lib.hs_server_get_stats(self._h, C.byref(st))
        return {k: int(getattr(st, k)) for k, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            _lib.hs_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
