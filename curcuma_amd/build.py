"""Build recipe for the engine's shared library and the CPU oracle.

``build_engine()`` compiles the C++ host code and the HIP kernels for gfx950
into ``curcuma_amd/lib/libgfnff_mi355x.so`` (in-tree, so it travels with the
repository).  ``build_oracle()`` compiles the test-only CPU oracle into
``oracle/_build/liboracle.so``.  Both always recompile.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
OBJDIR = PKG / "_build"
LIB = LIBDIR / "libgfnff_mi355x.so"
ORACLE_SRC = ROOT / "oracle" / "gfnff_oracle.c"
ORACLE_LIB = ROOT / "oracle" / "_build" / "liboracle.so"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or f"{ROCM}/bin/hipcc"
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
# packed FP32 instructions stay disabled in every device compile
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]

HOST_SOURCES = ["capi.cpp", "host/gfnff_topology.cpp", "host/alpb.cpp"]
DEVICE_SOURCES = ["device/engine.hip"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)


def build_engine(verbose: bool = False) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    jobs = []
    objs = []
    for src in HOST_SOURCES:
        obj = OBJDIR / (src.replace("/", "_") + ".o")
        jobs.append([HIPCC, *COMMON, "-x", "c++", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in DEVICE_SOURCES:
        obj = OBJDIR / (src.replace("/", "_") + ".o")
        jobs.append([HIPCC, *COMMON, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose: bool = False) -> Path:
    ORACLE_LIB.parent.mkdir(parents=True, exist_ok=True)
    _run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", str(ORACLE_SRC), "-o",
          str(ORACLE_LIB), "-lm"])
    if verbose:
        print("built", ORACLE_LIB)
    return ORACLE_LIB


def build_all(verbose: bool = False) -> None:
    build_engine(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
