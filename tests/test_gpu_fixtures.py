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
