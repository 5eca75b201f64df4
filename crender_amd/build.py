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


# Per-file flags.  The integrator's shading, camera and sampling code is compiled without
# FP contraction so that every product and sum rounds as written, like the CPU restatement
# it is checked against (the traversal keeps its explicit fmaf node test).
PER_FILE = {
    "integrator.hip": ["-ffp-contract=off"],
}


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    _run([HIPCC, *COMMON, *PER_FILE.get(src.name, []), "-c", str(src), "-o", str(obj)])
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


REFERENCE = Path("/root/reference")


def build_oracle(verbose: bool = True) -> Path:
    """Build the CPU oracle (test infrastructure) with its Makefile, and, where the reference
    tree is present, oracle/_ref/ from the reference's own sources (its vendored
    tinyobjloader)."""
    odir = ROOT / "oracle"
    make = shutil.which("make") or "make"
    _run([make, "-C", str(odir), "-s"])
    lib = odir / "liboracle.so"
    if verbose:
        print(f"built {lib}")
    if (REFERENCE / "external" / "tinyobj" / "tinobj.h").exists():
        _run([make, "-C", str(odir), "-s", "ref", f"REFERENCE={REFERENCE}"])
        if verbose:
            print(f"built {odir / '_ref' / 'tinyobj_dump'} from the reference's tinobj.h")
    return lib


if __name__ == "__main__":
    build()
    if "--oracle" in sys.argv:
        build_oracle()
