"""Builds libcmag_hip.so (gfx950) in-tree and, for the tests, the CPU oracle.

    python -m cmag_amd.build            # library + oracle
    python -m cmag_amd.build --lib      # library only

Every HIP translation unit is compiled with hipcc for gfx950 only, with packed
FP32 instructions disabled (-packed-fp32-ops), then linked into
cmag_amd/libcmag_hip.so.  Objects go to cmag_amd/_build/ (git-ignored).
"""
from __future__ import annotations

import argparse
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
LIB = PKG / "libcmag_hip.so"
ARCH = "gfx950"

SOURCES = [
    "params.cpp",
    "init.cpp",
    "system.cpp",
    "grid.hip",
    "sort.hip",
    "sph.hip",
    "integrate.hip",
    "batch.hip",
    "adjoint.hip",
    "render.hip",
]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libcmag_hip.so")
    return path


def _compile(src: Path, obj: Path) -> tuple[Path, str]:
    cmd = [hipcc(), *HIPCC_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{proc.stderr}")
    return obj, proc.stderr


def build_library(jobs: int | None = None, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, CSRC / s, BUILD / (Path(s).stem + ".o")) for s in SOURCES]
        for f in futs:
            obj, warn = f.result()
            objs.append(obj)
            if verbose and warn.strip():
                print(warn, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr}")
    os.replace(tmp, LIB)
    return LIB


EXAMPLES = ["poiseuille_report", "dambreak_fronts", "dambreak_dump", "dambreak_render"]


def build_examples() -> list[Path]:
    """C++ programs written against include/cmag/systems.hpp, linked to the library."""
    out = []
    for name in EXAMPLES:
        src = PKG / "examples" / f"{name}.cpp"
        exe = PKG / "examples" / name
        cmd = [hipcc(), "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
               f"-L{PKG}", "-lcmag_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/.."]
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError(f"example {name} failed to build:\n{proc.stderr}")
        out.append(exe)
    return out


def build_oracle() -> Path:
    """The oracle is test infrastructure; built beside the library for the tests."""
    odir = ROOT / "oracle"
    # the AVX2 library the tests load, and the AVX-512 one the CPU baseline
    # uses on hosts that have it
    for extra in ([], ["OUT=liboracle_v4.so", "MARCH=x86-64-v4"]):
        proc = subprocess.run(["make", "-s", "-C", str(odir)] + extra, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError(f"oracle build failed:\n{proc.stdout}\n{proc.stderr}")
    return odir / "liboracle.so"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("-v", action="store_true")
    args = ap.parse_args()
    print(build_library(verbose=args.v))
    if not args.lib:
        print(build_examples())
        print(build_oracle())


if __name__ == "__main__":
    main()
