"""Builds libhyset_amd.so (HIP kernels for gfx950 + C ABI) in-tree.

Every HIP source is compiled with hipcc for gfx950 only; host-only C++
sources are compiled by hipcc as plain C++.  The result lands next to this
file so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libhyset_amd.so"
DRIVER = HERE / "hyset_join"

HIP_SOURCES = ["index.hip", "fgss.hip", "bitmap.hip", "verify.hip", "mlp.hip", "cost.hip", "noalloc.hip", "tokindex.hip"]
CPP_SOURCES = ["api.cpp", "batch.cpp", "serve.cpp", "host_join.cpp", "bounds.cpp", "datagen.cpp", "host_tools.cpp",
               "preprocess.cpp", "coprocess.cpp", "partition.cpp"]
ARCH = "gfx950"


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def _common_flags() -> list[str]:
    return ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
            f"-I{ROOT / 'include'}", f"-I{CSRC}", "-ffp-contract=off"]


def _compile(src: Path, obj: Path) -> None:
    if src.suffix == ".hip":
        cmd = [hipcc(), "-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
               "-Xclang", "-packed-fp32-ops", *_common_flags(), "-c", str(src), "-o", str(obj)]
    else:
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        cmd = [hipcc(), "-x", "c++", *_common_flags(), "-D__HIP_PLATFORM_AMD__", f"-I{rocm}/include",
               "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> Path:
    """Compiles every source and links the shared library; returns its path."""
    BUILD.mkdir(exist_ok=True)
    jobs = []
    for name in HIP_SOURCES + CPP_SOURCES:
        src = CSRC / name
        jobs.append((src, BUILD / (src.stem + src.suffix.replace(".", "_") + ".o")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        futs = [ex.submit(_compile, s, o) for s, o in jobs]
        for f in futs:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *[str(o) for _, o in jobs], "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    # C++ command-line driver over the handler layer, linked to the library
    drv_tmp = DRIVER.with_suffix(".tmp")
    cmd = ["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(HERE / "drivers" / "hyset_join.cpp"),
           "-o", str(drv_tmp), f"-L{HERE}", "-lhyset_amd", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(drv_tmp, DRIVER)
    if verbose:
        print(f"built {LIB} and {DRIVER}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
