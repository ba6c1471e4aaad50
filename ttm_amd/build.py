"""Build the gfx950 library ``ttm_amd/lib/libttm_amd.so`` in-tree with hipcc.

Usage: ``python -m ttm_amd.build`` (or ``ttm_amd.build.build()``).  Every
translation unit under ``ttm_amd/csrc`` is compiled for ``--offload-arch=gfx950``
and linked into one shared library whose extern "C" surface is
``include/ttm_amd.h``.
"""
from __future__ import annotations

import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
BUILD = PKG / "_build"
LIB = LIBDIR / "libttm_amd.so"

ARCH = "gfx950"
# Packed FP32 instructions are disabled for every device pass.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
          "-Wno-unused-command-line-argument", "-Wno-unused-result",
          f"-I{ROOT / 'include'}", f"-I{CSRC}"]


# Per-file flags: the host executor sums without contraction into FMAs, so
# its results do not depend on the host CPU's instruction set.
EXTRA = {"ttm_cpu.cpp": ["-ffp-contract=off"]}


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _compile(hipcc: str, src: pathlib.Path) -> pathlib.Path:
    obj = BUILD / (src.name + ".o")
    cmd = [hipcc, *COMMON, *DEVICE_FLAGS, *EXTRA.get(src.name, []), "-c", str(src), "-o", str(obj)]
    if src.suffix == ".hip":
        cmd[1:1] = ["-x", "hip"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> pathlib.Path:
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    sources = sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.cpp")])
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(lambda s: _compile(hipcc, s), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB.relative_to(ROOT)} from {len(sources)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
