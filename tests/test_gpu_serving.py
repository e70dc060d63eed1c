"""Serving: every request's answer is the oracle's R-S join of its query
sets against the index at its threshold, the same alone and in the mix;
the byte budget defers or refuses requests without changing answers."""
import random

import numpy as np
import pytest

import hyset_amd as H
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(ctx):
    idx, q = H.generate_queries(H.gen_params(30000, median=12, sigma=0.6, lo=2, hi=120, universe=30000, zipf=1.1,
                                             seed=1),
                                H.gen_params(6000, median=12, sigma=0.6, lo=2, hi=120, universe=30000, zipf=1.1,
                                             seed=2))
    return idx, q, ctx.upload(idx)


def requests(q, k, seed):
    rnd = random.Random(seed)
    out, pos = [], 0
    for _ in range(k):
        n = max(1, int(round(np.exp(rnd.uniform(0, np.log(400))))))
        if pos + n > q.n:
            pos = 0
        sets = [q.set(i).tolist() for i in range(pos, pos + n)]
        pos += n
        out.append((H.SetCollection.from_sets(sets), rnd.choice([0.6, 0.7, 0.8, 0.9])))
    return out


def triples(r, s, ov):
    r, s, ov = H.sort_pairs(r, s, ov)
    return list(zip(r.tolist(), s.tolist(), ov.tolist()))


def oracle_answer(qc, idx, t, sim="jaccard"):
    res = O.rs_join(qc, idx, t, sim)
    return list(zip(res.r.tolist(), res.s.tolist(), res.overlap.tolist()))


def test_alone_equals_mix_equals_oracle(served):
    idx, q, d = served
    reqs = requests(q, 40, 7)
    with H.Server(d, slots=4) as srv:
        alone = []
        for qc, t in reqs:
            r, s, ov, _ = srv.wait_pairs(srv.submit(qc, t))
            alone.append(triples(r, s, ov))
        ids = [srv.submit(qc, t) for qc, t in reqs]
        mix = [triples(*srv.wait_pairs(i)[:3]) for i in ids]
        st = srv.stats()
    assert mix == alone
    total = 0
    for (qc, t), got in zip(reqs, alone):
        assert got == oracle_answer(qc, idx, t)
        total += len(got)
    assert total > 100
    assert st["batches"] < 80  # requests of the mix were batched together


def test_cosine_server(ctx, served):
    idx, q, d = served
    reqs = requests(q, 12, 3)
    with H.Server(d, sim="cosine", slots=2) as srv:
        ids = [srv.submit(qc, t) for qc, t in reqs]
        got = [triples(*srv.wait_pairs(i)[:3]) for i in ids]
    for (qc, t), g in zip(reqs, got):
        assert g == oracle_answer(qc, idx, t, "cosine")


def test_budget_defers_and_refuses(served):
    idx, q, d = served
    reqs = requests(q, 16, 11)
    big = max(reqs, key=lambda x: x[0].n)
    with H.Server(d, slots=1) as probe_srv:
        sig_bytes = probe_srv.stats()["signature_bits"] // 8
    assert sig_bytes == 8  # 64 bits for sets of mean size 12
    need = 4 * int(big[0].tokens.shape[0]) + (109 + sig_bytes) * big[0].n + 4096
    with H.Server(d, slots=4, hbm_budget_bytes=need) as srv:
        ids = [srv.submit(qc, t) for qc, t in reqs]
        got = [triples(*srv.wait_pairs(i)[:3]) for i in ids]
        st = srv.stats()
    for (qc, t), g in zip(reqs, got):
        assert g == oracle_answer(qc, idx, t)
    # the bytes in flight never exceed the budget; whether a request found
    # the budget taken and waited (st["deferred"]) depends on whether a
    # batch was still on the device when the next submit came, a race the
    # test does not control
    assert st["max_bytes_in_flight"] <= need and st["batches"] >= 1
    with H.Server(d, slots=2, hbm_budget_bytes=1024) as srv:
        rid = srv.submit(reqs[0][0], 0.8)
        with pytest.raises(H.HysetError) as e:
            srv.wait_pairs(rid)
        assert e.value.code == H.HS_E_BUSY
        assert srv.stats()["rejected"] == 1


def test_threshold_below_minimum_rejected(served):
    idx, q, d = served
    with H.Server(d, min_threshold=0.7) as srv:
        with pytest.raises(H.HysetError):
            srv.submit(requests(q, 1, 1)[0][0], 0.6)


@pytest.mark.parametrize("max_batch_queries", [65536, 2])
def test_result_larger_than_first_fetch(ctx, max_batch_queries):
    """A batch's result rows come back with its counters up to as many
    triples as the batch has queries (at least 1024); a batch with more
    pairs than that fetches the rest in a second copy.  Two queries that
    equal 3000 indexed sets give 6000 pairs: the second copy runs and the
    pairs are still the oracle's.  With a 2-query batch limit the slots are
    sized for 8 result triples, so the device rows regrow (and the batch
    runs again) and the host rows regrow and take every triple anew."""
    base = list(range(1, 11))
    sets = [base] * 3000 + [list(range(20, 20 + k)) for k in range(2, 40)]
    idx = H.SetCollection.from_sets(sorted(sets, key=lambda s: (len(s), s)))
    d = ctx.upload(idx)
    qc = H.SetCollection.from_sets([base, base])
    with H.Server(d, slots=2, max_batch_queries=max_batch_queries) as srv:
        for t in (0.8, 0.6):
            r, s, ov, _ = srv.wait_pairs(srv.submit(qc, t))
            got = triples(r, s, ov)
            assert got == oracle_answer(qc, idx, t)
            assert len(got) == 6000
