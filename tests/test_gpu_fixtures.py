"""GPU results against the committed golden fixtures, through the Python
view of the C ABI and through the C++ handler-layer driver."""
import subprocess
from pathlib import Path

import pytest

import hyset_amd as H
from tests.golden import make_fixtures as F

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
DRIVER = Path(H.__file__).resolve().parent / "hyset_join"


def triples(r, s, ov):
    r, s, ov = H.sort_pairs(r, s, ov)
    return list(zip(r.tolist(), s.tolist(), ov.tolist()))


@pytest.mark.parametrize("algo", ["fgss", "bitmap", "hybrid"])
def test_special_fixture(ctx, algo):
    c = H.read_collection_text(GOLD / "special.txt")
    d = ctx.upload(c)
    for sim, t in F.SPECIAL_CASES:
        if algo != "fgss" and sim != "jaccard":
            continue
        ctx.set_threshold(t, sim)
        assert triples(*ctx.join(d, algo=algo)) == F.read_pairs(GOLD / f"special_t{t}_{sim}.pairs"), (sim, t)


@pytest.mark.parametrize("algo", ["fgss", "bitmap", "hybrid"])
def test_neardup_fixture(ctx, algo):
    c = F.neardup()
    d = ctx.upload(c)
    for t in F.NEARDUP_T:
        ctx.set_threshold(t)
        tune = H.Tune(bitmap_bits=128) if algo == "bitmap" else None
        assert triples(*ctx.join(d, algo=algo, tune=tune)) == F.read_pairs(GOLD / f"neardup_t{t}.pairs"), t


def test_config1_fixture(ctx):
    c = F.config1()
    d = ctx.upload(c)
    for t in F.CONFIG1_T:
        ctx.set_threshold(t)
        assert triples(*ctx.join(d, algo="fgss")) == F.read_pairs(GOLD / f"config1_t{t}.pairs"), t


@pytest.mark.parametrize("algo,extra", [("fgss", []), ("bitmap", ["--bitmap", "256"]),
                                        ("hybrid", ["--scenario", "1"]), ("hybrid", ["--scenario", "2"])])
def test_cpp_driver(tmp_path, algo, extra):
    col = tmp_path / "neardup.txt"
    H.write_collection_text(col, F.neardup())
    out = tmp_path / "pairs.txt"
    for t in (0.7, 0.9):
        res = subprocess.run([str(DRIVER), "--input", str(col), "--threshold", str(t), "--block", "1000",
                              "--algorithm", algo, "--output", str(out), "--overlap", *extra],
                             capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stderr
        got = sorted(tuple(int(x) for x in line.split()) for line in out.read_text().splitlines())
        assert got == F.read_pairs(GOLD / f"neardup_t{t}.pairs")
        res = subprocess.run([str(DRIVER), "--input", str(col), "--threshold", str(t), "--algorithm", algo, *extra],
                             capture_output=True, text=True, timeout=120)
        assert res.returncode == 0 and f"Result: {len(got)}" in res.stdout


@pytest.mark.parametrize("args", [["--algorithm", "ppjoin"], ["--algorithm", "allpairs", "--scenario", "2"],
                                  ["--technique", "cpu", "--algorithm", "allpairs", "--threads", "4"],
                                  ["--technique", "cpu", "--algorithm", "ppjoin", "--threads", "1"]])
def test_cpp_driver_cpu_algorithms(tmp_path, args):
    """The reference's --algorithm allpairs|ppjoin: alone it selects the
    hybrid technique with that host filter (src/hybrid.cu); with --technique
    cpu the join runs on host threads (src/cpu_single.cu, src/cpu_multi.cu).
    Pairs and overlaps as the golden fixture."""
    col = tmp_path / "neardup.txt"
    H.write_collection_text(col, F.neardup())
    out = tmp_path / "pairs.txt"
    for t in (0.7, 0.9):
        res = subprocess.run([str(DRIVER), "--input", str(col), "--threshold", str(t), "--block", "1000",
                              "--output", str(out), "--overlap", *args], capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stderr
        got = sorted(tuple(int(x) for x in line.split()) for line in out.read_text().splitlines())
        assert got == F.read_pairs(GOLD / f"neardup_t{t}.pairs"), (args, t)
    res = subprocess.run([str(DRIVER), "--input", str(col), "--algorithm", "quicksort"], capture_output=True,
                         text=True, timeout=60)
    assert res.returncode != 0


def _coefficient_file(path, t):
    """A coefficient table in the reference's format (src/cost_model.cu:296-323):
    the CPU cheapest for small partition pairs, fgss for large ones, bitmap
    and hybrid never (their intercepts are large)."""
    import json
    key = f"{t:.2g}"
    base = dict(Intercept=0.0, set_size_std=0.0, avg_set_size=0.0, token_freq=0.0, dataset_size=0.0,
                dataset_size_square=0.0, universe_size=0.0)
    tab = {"cpu": dict(base, dataset_size=1e-3, index_size=0.0),
           "fgssjoin": dict(base, Intercept=1.5, index_size=0.0),
           "hybrid": dict(base, Intercept=1e6, index_size=0.0),
           "bitmap": dict(base, Intercept=1e6, bitmap=0.0)}
    path.write_text(json.dumps({k: {key: v, "0.5": v} for k, v in tab.items()}))


@pytest.mark.parametrize("args", [["--technique", "queue", "--cpu-threads", "2"],
                                  ["--technique", "queue", "--cooperative", "--cpu-threads", "2"],
                                  ["--technique", "queue", "--bitmap", "128", "--cpu-threads", "2"],
                                  ["--technique", "dichotomy", "--policy", "cpu,gpu", "--split", "40"],
                                  ["--technique", "dichotomy", "--policy", "gpu,cpu", "--split", "40"],
                                  ["--technique", "dichotomy", "--policy", "cpu,cpu-gpu", "--split", "70"],
                                  ["--technique", "dichotomy", "--policy", "cpu-gpu,cpu", "--split", "20",
                                   "--algorithm", "ppjoin"],
                                  ["--technique", "cost", "--partitions", "3", "--cpu-threads", "2"]])
def test_cpp_driver_cpu_and_device_together(tmp_path, args):
    """The reference's queue, dichotomy and cost_model drivers: CPU threads
    and the device on one join, pairs and overlaps as the golden fixture."""
    col = tmp_path / "neardup.txt"
    H.write_collection_text(col, F.neardup())
    out = tmp_path / "pairs.txt"
    coef = tmp_path / "coefficients.json"
    extra = ["--coefficients", str(coef)] if "cost" in args else []
    for t in (0.7, 0.9):
        _coefficient_file(coef, t)
        res = subprocess.run([str(DRIVER), "--input", str(col), "--threshold", str(t), "--block", "100",
                              "--output", str(out), "--overlap", *args, *extra], capture_output=True, text=True,
                             timeout=120)
        assert res.returncode == 0, res.stderr
        got = sorted(tuple(int(x) for x in line.split()) for line in out.read_text().splitlines())
        assert got == F.read_pairs(GOLD / f"neardup_t{t}.pairs"), (args, t)
        res = subprocess.run([str(DRIVER), "--input", str(col), "--threshold", str(t), "--block", "100", *args,
                              *extra], capture_output=True, text=True, timeout=120)
        assert res.returncode == 0 and f"Result: {len(got)}" in res.stdout, res.stdout + res.stderr
        if "cost" in args:
            assert "cpu:" in res.stdout and "fgssjoin:" in res.stdout  # both chosen for some pairs
    for bad in (["--technique", "dichotomy", "--policy", "gpu,cpu-gpu"], ["--technique", "dichotomy", "--policy",
                "cpu,cpu"], ["--technique", "dichotomy", "--split", "120"], ["--technique", "cost"]):
        res = subprocess.run([str(DRIVER), "--input", str(col), *bad], capture_output=True, text=True, timeout=60)
        assert res.returncode != 0, bad


def _names(n, seed):
    """Strings of 8-20 letters, 30% of them copies of earlier ones with 1-2
    character edits (the string-matching use of the join)."""
    import random
    rnd = random.Random(seed)
    abc = "abcdefghijklmnopqrstuvwxyz"
    out = []
    for _ in range(n):
        if out and rnd.random() < 0.3:
            s = list(rnd.choice(out))
            for _ in range(rnd.randint(1, 2)):
                k = rnd.randrange(len(s))
                s[k] = rnd.choice(abc)
            out.append("".join(s))
        else:
            out.append("".join(rnd.choice(abc) for _ in range(rnd.randint(8, 20))))
    return out


@pytest.mark.parametrize("algo", ["fgss", "bitmap"])
def test_prepared_qgram_join(ctx, algo):
    """Raw strings prepared into 3-gram sets join on the device exactly as in
    the oracle."""
    import oracle as O
    c, rec = H.preprocess_text("\n".join(_names(4000, 5)), mode="qgrams", q=3)
    d = ctx.upload(c)
    try:
        for t in (0.6, 0.8):
            ctx.set_threshold(t)
            want = O.self_join(c, t)
            assert want.count > 20
            got = triples(*ctx.join(d, algo=algo))
            assert got == list(zip(want.r.tolist(), want.s.tolist(), want.overlap.tolist())), t
    finally:
        d.free()


def test_cpp_driver_prepare_and_binary(tmp_path):
    """The driver prepares raw text (--prepare), saves and reloads the binary
    cache (--save-binary, --binary), and maps set ids back to input lines."""
    import oracle as O
    names = _names(3000, 6)
    raw = tmp_path / "names.txt"
    raw.write_text("\n".join(names) + "\n")
    c, rec = H.preprocess_text("\n".join(names), mode="qgrams", q=3)
    want = O.self_join(c, 0.7)
    want_t = sorted(zip(want.r.tolist(), want.s.tolist()))
    out, cache = tmp_path / "pairs.txt", tmp_path / "names.bin"
    res = subprocess.run([str(DRIVER), "--input", str(raw), "--prepare", "qgrams", "--q", "3", "--threshold", "0.7",
                          "--save-binary", str(cache), "--output", str(out)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    got = sorted(tuple(int(x) for x in l.split()[:2]) for l in out.read_text().splitlines() if l.strip())
    assert got == want_t
    recs = [tuple(int(x) for x in l.split()) for l in (tmp_path / "pairs.txt.records").read_text().splitlines()]
    assert recs == list(enumerate(rec.tolist()))
    out2 = tmp_path / "pairs2.txt"
    res = subprocess.run([str(DRIVER), "--input", str(cache), "--binary", "--threshold", "0.7", "--output", str(out2)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    got2 = sorted(tuple(int(x) for x in l.split()[:2]) for l in out2.read_text().splitlines() if l.strip())
    assert got2 == want_t
