"""Builds the crt HIP library (libcrt.so) for gfx950 in-tree.

Every .hip/.cpp file under crender_amd/csrc is compiled with hipcc to an object in
crender_amd/_build/ and linked into crender_amd/libcrt.so; crender_amd/examples/*.cpp
are compiled against it.  Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libcrt.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    _run([HIPCC, *COMMON, "-c", str(src), "-o", str(obj)])
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile all HIP sources and link libcrt.so; returns the library path."""
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread", "-lz"])
    os.replace(tmp, LIB)
    for ex_src in sorted((PKG / "examples").glob("*.cpp")):
        exe = PKG / "examples" / ex_src.stem
        _run([HIPCC, "-O2", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{PKG / 'host'}", str(ex_src),
              "-o", str(exe), f"-L{PKG}", "-lcrt", f"-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print(f"built {LIB} ({LIB.stat().st_size // 1024} KiB) from {len(srcs)} sources")
    return LIB


def build_oracle(verbose: bool = True) -> Path:
    """Build the CPU oracle (test infrastructure) with its Makefile."""
    odir = ROOT / "oracle"
    make = shutil.which("make") or "make"
    _run([make, "-C", str(odir), "-s"])
    lib = odir / "liboracle.so"
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    build()
    if "--oracle" in sys.argv:
        build_oracle()
