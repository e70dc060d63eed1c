"""hs_join_noalloc, the join that allocates nothing: through the C ABI, its
triples equal hs_join's (and the oracle's) on consecutive calls of
decreasing size, it keeps the device's free memory unchanged, reports
overflow with the count needed and refuses joins beyond its scratch."""
import ctypes as C

import numpy as np
import pytest

import hyset_amd as H
import oracle as O
from hyset_amd import workloads as W

pytestmark = pytest.mark.gpu


_hip = None


def hip():
    """The HIP runtime the library itself links (already loaded with it)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DevArrays:
    """Three uint32 device arrays owned by the caller (hipMalloc)."""

    def __init__(self, capacity):
        self.capacity = int(capacity)
        self.p = []
        for _ in range(3):
            p = C.c_void_p()
            assert hip().hipMalloc(C.byref(p), max(self.capacity, 1) * 4) == 0
            self.p.append(p)

    def ptrs(self):
        return dict(r=self.p[0].value, s=self.p[1].value, overlap=self.p[2].value, capacity=self.capacity)

    def triples(self, n):
        out = []
        for p in self.p:
            a = np.empty(max(n, 1), dtype=np.uint32)
            if n:
                assert hip().hipMemcpy(a.ctypes.data, p, n * 4, 2) == 0  # device to host
            out.append(a[:n])
        return H.sort_pairs(*out)

    def __del__(self):
        for p in self.p:
            hip().hipFree(p)


def _same(a, b, what):
    for x, y in zip(a, b):
        assert np.array_equal(x, y), what


def _free_bytes():
    free, total = C.c_size_t(), C.c_size_t()
    assert hip().hipDeviceSynchronize() == 0
    assert hip().hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_noalloc_equals_hs_join_on_two_calls(ctx):
    """The 1M-set self-join at Jaccard 0.8, then a smaller 50K-set join: the
    same triples as hs_join (and the oracle), no change in the device's free
    memory across the second call of each."""
    big = W.headline()
    small = H.generate_self(H.gen_params(50_000, median=14, sigma=0.5, lo=2, hi=120, universe=40_000, zipf=1.0,
                                         seed=9))
    ctx.set_threshold(0.8)
    for c, name in ((big, "1M"), (small, "50K")):
        d = ctx.upload(c)
        try:
            want = H.sort_pairs(*ctx.join(d, algo="fgss"))
            st_join = ctx.stats()
            out = DevArrays(want[0].shape[0] + 64)
            n = ctx.join_noalloc(d, **out.ptrs())
            assert n == want[0].shape[0], name
            _same(out.triples(n), want, f"noalloc {name}")
            st = ctx.stats()
            for k in ("posting_visits", "candidates", "signature_tests", "index_entries", "pairs"):
                assert st[k] == st_join[k], (name, k)
            # a second call allocates nothing: the device's free memory is unchanged
            before = _free_bytes()
            assert ctx.join_noalloc(d, **out.ptrs()) == n
            assert _free_bytes() == before, name
            _same(out.triples(n), want, f"noalloc {name} again")
            o = O.self_join(c, 0.8)
            _same(out.triples(n), (o.r, o.s, o.overlap), f"noalloc {name} vs oracle")
        finally:
            d.free()


def test_noalloc_ranges_rs_and_knobs(ctx):
    """Block pairs, an R-S join with cosine, fixed tiles, other signature
    widths and no signature test: as hs_join each time."""
    c = H.generate_self(H.gen_params(20_000, median=10, sigma=0.6, lo=2, hi=200, universe=20_000, zipf=1.1, seed=4))
    R, S = H.generate_rs(H.gen_params(3000, size_dist=H.SIZE_UNIFORM, lo=20, hi=200, universe=5000, zipf=0.8, seed=3),
                         H.gen_params(3000, size_dist=H.SIZE_UNIFORM, lo=20, hi=200, universe=5000, zipf=0.8, seed=4),
                         dup_frac=0.3)
    d, dR, dS = ctx.upload(c), ctx.upload(R), ctx.upload(S)
    out = DevArrays(1 << 18)
    try:
        ctx.set_threshold(0.7)
        for ir, pr in (((0, 9999), (5000, 19999)), ((12000, 19999), (12000, 19999)), ((0, 0), (0, 19999))):
            want = H.sort_pairs(*ctx.join(d, algo="fgss", indexed_range=ir, probe_range=pr))
            n = ctx.join_noalloc(d, indexed_range=ir, probe_range=pr, **out.ptrs())
            _same(out.triples(n), want, f"block {ir} {pr}")
        for tune in (H.Tune(probes_per_tile=37), H.Tune(signature_bits=256), H.Tune(flags=H.TUNE_NO_SIGNATURE)):
            want = H.sort_pairs(*ctx.join(d, algo="fgss", tune=tune))
            n = ctx.join_noalloc(d, tune=tune, **out.ptrs())
            _same(out.triples(n), want, f"tune {tune.probes_per_tile} {tune.signature_bits} {tune.flags}")
        ctx.set_threshold(0.8, "cosine")
        want = H.sort_pairs(*ctx.join(dS, dR, algo="fgss"))
        assert want[0].shape[0] > 100
        n = ctx.join_noalloc(dS, dR, **out.ptrs())
        _same(out.triples(n), want, "R-S cosine")
        o = O.rs_join(R, S, 0.8, "cosine")
        _same(out.triples(n), (o.r, o.s, o.overlap), "R-S cosine vs oracle")
    finally:
        ctx.set_threshold(0.8)
        for x in (d, dR, dS):
            x.free()


def test_noalloc_overflow_and_limit(ctx):
    """Fewer slots than pairs: HS_E_OVERFLOW with the count needed, then the
    full result with room for it.  A join whose index does not fit the
    scratch array (the 1M-set join at Jaccard 0.6, 3.6M index entries) is
    refused before any device work."""
    c = W.headline()
    d = ctx.upload(c)
    try:
        ctx.set_threshold(0.8)
        want = H.sort_pairs(*ctx.join(d, algo="fgss"))
        k = want[0].shape[0]
        small = DevArrays(k // 3)
        with pytest.raises(H.HysetError) as e:
            ctx.join_noalloc(d, **small.ptrs())
        assert e.value.code == H.HS_E_OVERFLOW and e.value.count == k
        with pytest.raises(H.HysetError) as e:
            ctx.join_noalloc(d, r=0, s=0, overlap=0, capacity=0)
        assert e.value.code == H.HS_E_OVERFLOW and e.value.count == k
        full = DevArrays(k)
        assert ctx.join_noalloc(d, **full.ptrs()) == k
        _same(full.triples(k), want, "after overflow")
        ctx.set_threshold(0.6)
        with pytest.raises(H.HysetError) as e:
            ctx.join_noalloc(d, **full.ptrs())
        assert e.value.code == H.HS_E_UNSUPPORTED and "scratch" in str(e.value)
        # and the next call that fits still works
        ctx.set_threshold(0.9)
        want9 = H.sort_pairs(*ctx.join(d, algo="fgss"))
        big = DevArrays(want9[0].shape[0])
        n = ctx.join_noalloc(d, **big.ptrs())
        _same(big.triples(n), want9, "t=0.9 after refusal")
    finally:
        ctx.set_threshold(0.8)
        d.free()
