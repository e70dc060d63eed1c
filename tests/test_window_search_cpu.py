"""The posting-list window search of k_fgss and k_item_slices, compiled for
the host and checked against std::lower_bound.

`lockstep_bounds` (hyset_amd/csrc/fgss.hip) finds, for K lists at once, the
slice of ids in [k_lo, k_hi) by reading aligned 16-byte quads of ids and
tightening both bounds from every id it reads.  The device function is taken
from the source text as it stands (between the NO_QUAD constant and the
probe-window helper), built with g++ under host stand-ins for the few HIP
names it uses, and run on random sorted lists laid end to end in one buffer
(as the index lays them out) that is exactly as long as the index's id array
with its IDS_PAD padding, under AddressSanitizer, so a quad read past that
padding fails the test.  Cases: empty lists, lists of 1-5 ids (every
alignment), lists up to 3000 ids with many repeats, keys below, inside and
above a list, k_lo == k_hi, and 2- and 4-quad rounds with 1-4 lists in
lockstep.
"""
from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "hyset_amd" / "csrc" / "fgss.hip"
HDR = ROOT / "hyset_amd" / "csrc" / "hs_internal.h"

HARNESS_HEAD = r"""
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#define __device__
#define __forceinline__ inline
#define __restrict__
struct uint4 { uint32_t x, y, z, w; };
static inline uint4 make_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return uint4{a, b, c, d}; }
using std::max;
using std::min;
"""

HARNESS_MAIN = r"""
template <uint32_t K, uint32_t WQ>
static int run(std::mt19937& rng, int trials) {
  for (int trial = 0; trial < trials; ++trial) {
    std::vector<uint32_t> ids;
    uint32_t f[K], n[K], klo[K], khi[K];
    const int shape = rng() % 5;
    for (uint32_t k = 0; k < K; ++k) {
      const uint32_t len = shape == 0 ? rng() % 6 : shape == 1 ? rng() % 40 : shape == 2 ? rng() % 300
                         : shape == 3 ? rng() % 3000 : (rng() % 2 ? 0u : 1u + rng() % 4);
      f[k] = (uint32_t)ids.size();
      n[k] = len;
      const uint32_t range = 1 + rng() % (shape == 3 ? 50000 : 2000);
      std::vector<uint32_t> l(len);
      for (auto& x : l) x = rng() % range;
      std::sort(l.begin(), l.end());
      ids.insert(ids.end(), l.begin(), l.end());
      uint32_t a = rng() % (range + 2), b = rng() % (range + 2);
      if (a > b) std::swap(a, b);
      if (rng() % 8 == 0) b = a;
      klo[k] = a;
      khi[k] = b;
    }
    // the index's id array: the entries, then IDS_PAD ids of padding, and nothing more
    const size_t total = ids.size() + IDS_PAD;
    uint32_t* buf = static_cast<uint32_t*>(std::malloc(total * 4));
    std::copy(ids.begin(), ids.end(), buf);
    for (size_t i = ids.size(); i < total; ++i) buf[i] = 0xFFFFFFFFu - (uint32_t)i;
    uint32_t f_lo[K], n_lo[K], n_hi[K];
    for (uint32_t k = 0; k < K; ++k) {
      f_lo[k] = f[k];
      n_lo[k] = n[k];
      n_hi[k] = n[k];
    }
    lockstep_bounds<K, WQ>(buf, f_lo, n_lo, n_hi, klo, khi);
    std::free(buf);
    for (uint32_t k = 0; k < K; ++k) {
      auto b0 = ids.begin() + f[k], e0 = b0 + n[k];
      const uint32_t a = (uint32_t)(std::lower_bound(b0, e0, klo[k]) - b0);
      const uint32_t b = (uint32_t)(std::lower_bound(b0, e0, khi[k]) - b0);
      if (f_lo[k] != f[k] + a || n_hi[k] != b - a) {
        std::printf("K=%u WQ=%u trial %d list %u (n=%u, k_lo=%u, k_hi=%u): slice [%u, +%u), want [%u, +%u)\n", K, WQ,
                    trial, k, n[k], klo[k], khi[k], f_lo[k] - f[k], n_hi[k], a, b - a);
        return 1;
      }
    }
  }
  return 0;
}

int main() {
  std::mt19937 rng(12345);
  int bad = run<1, 2>(rng, 20000) + run<3, 2>(rng, 20000) + run<2, 4>(rng, 20000) + run<4, 4>(rng, 10000) +
            run<1, 4>(rng, 20000);
  if (bad) return 1;
  std::printf("ok\n");
  return 0;
}
"""


def _extract() -> str:
    src = SRC.read_text()
    start = src.index("constexpr uint32_t NO_QUAD")
    end = src.index("// Indexed ids [lo, hi) that probe p of size l joins:")
    body = src[start:end]
    # the record loader between the two needs the device index type: drop it
    i = body.index("// Index record e:")
    j = body.index("// Plans m quads")
    body = body[:i] + body[j:]
    pad = re.search(r"constexpr uint32_t IDS_PAD = (\d+);", HDR.read_text())
    assert pad, "IDS_PAD not found"
    return f"constexpr uint32_t IDS_PAD = {pad.group(1)};\n" + body


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_window_search_matches_lower_bound(tmp_path):
    cpp = tmp_path / "search.cpp"
    cpp.write_text(HARNESS_HEAD + _extract() + HARNESS_MAIN)
    exe = tmp_path / "search"
    build = subprocess.run(["g++", "-O1", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                            "-o", str(exe), str(cpp)], capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-4000:]
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)  # no progress: a hang
    assert res.returncode == 0, (res.stdout + res.stderr)[-4000:]
    assert res.stdout.strip().endswith("ok")
