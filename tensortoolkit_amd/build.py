"""Build the in-tree native libraries.

* ``tensortoolkit_amd/libqlten_amd.so`` -- every HIP unit under ``csrc/``,
  compiled for gfx950 with hipcc and linked into one shared library that
  exports the C ABI declared in ``include/qlten_amd.h``.
* ``oracle/liboracle.so`` -- the CPU restatement of the reference algorithm
  (test and baseline infrastructure only; built here so that one ``build()``
  call prepares everything the tests and the benchmark load).
* ``tests/cpp/bin/*`` -- C++ programs written against the host ``qlten``
  headers (the drop-in API), linked to ``libqlten_amd.so``.

Object files go to ``_build/`` (git-ignored).  Every call recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "tensortoolkit_amd"
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libqlten_amd.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
CPP_TEST_DIR = ROOT / "tests" / "cpp"
CPP_BIN = CPP_TEST_DIR / "bin"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _compile_hip(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def build_hip_library(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(sources))) as ex:
        objs = list(ex.map(_compile_hip, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-Wall",
          "-o", str(tmp), str(ORACLE_DIR / "oracle.c"), "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_cpp_tests() -> list[Path]:
    outs = []
    srcs = sorted(CPP_TEST_DIR.glob("*.cc")) if CPP_TEST_DIR.exists() else []
    if not srcs:
        return outs
    CPP_BIN.mkdir(exist_ok=True)

    def one(src: Path) -> Path:
        exe = CPP_BIN / src.stem
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}",
              "-o", str(exe), str(src), f"-L{PKG}", "-lqlten_amd",
              f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../../tensortoolkit_amd"])
        return exe

    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        outs = list(ex.map(one, srcs))
    return outs


def build_all() -> None:
    build_hip_library()
    build_oracle()
    build_cpp_tests()


if __name__ == "__main__":
    build_all()
    print("built", LIB, ORACLE_LIB)
