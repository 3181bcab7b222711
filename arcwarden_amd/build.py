"""Build recipe for libarcwarden_hip.so (gfx950) -- the product library.

Compiles every ``csrc/*.hip`` with hipcc for gfx950 (and the host-only
``csrc/*.cpp`` with g++) and links them with rocFFT
into ``arcwarden_amd/libarcwarden_hip.so`` (in-tree, so it travels with the
repository snapshot).  Usage: ``python -m arcwarden_amd.build`` or
``arcwarden_amd.build.build()``.
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
OBJ = PKG / "_build"
LIB = PKG / "libarcwarden_hip.so"

ARCH = os.environ.get("AW_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++20",
    # every fused multiply-add in the kernels is an explicit fmaf; nothing else
    # may be contracted, so the FP32 sequence matches the host restatement
    "-ffp-contract=off",
    # no packed-FP32 VALU forms
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


# host-only sources (no device code): plain g++
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++20", "-Wall", f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + (".o" if src.suffix == ".hip" else ".cpp.o"))
    if src.suffix == ".hip":
        cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [os.environ.get("CXX", "g++"), *CXX_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    workers = min(len(sources), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs),
           "-L/opt/rocm/lib", "-lrocfft", "-Wl,-rpath,/opt/rocm/lib", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB.relative_to(ROOT)} from {len(sources)} sources", flush=True)
    for src, exe, extra in PROGRAMS:
        _link_program(src, exe, extra)
        if verbose:
            print(f"built {exe.relative_to(ROOT)}", flush=True)
    return LIB


# C++ programs on the facade (include/pic) linked against the library; the
# run-path is relative to the executable so the tree can move.
# (source, executable, extra flags): arcsim and run_deck are one runner with
# two command lines.
PROGRAMS = [
    (PKG / "tools" / "arcsim.cpp", PKG / "tools" / "arcsim", []),
    (PKG / "tools" / "arcsim.cpp", PKG / "tools" / "run_deck", ["-DARC_RUNNER_RUN_DECK=1"]),
    (ROOT / "tests" / "cpp" / "facade_check.cpp", ROOT / "tests" / "cpp" / "facade_check", []),
]


def _link_program(src: Path, exe: Path, extra: list) -> None:
    rel = os.path.relpath(PKG, exe.parent)
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++20", "-Wall", *extra,
           f"-I{ROOT / 'include'}", f"-I{PKG / 'include'}", str(src), "-o", str(exe),
           f"-L{PKG}", "-larcwarden_hip", f"-Wl,-rpath,$ORIGIN/{rel}",
           "-Wl,-rpath,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build failed for {src.name}:\n{res.stdout}\n{res.stderr}")


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
