"""Build the native libraries in-tree.

* ``ma57_amd/libsymla_hip.so`` -- the product: gfx950 HIP kernels, the CPU
  symbolic analysis and the C-ABI declared in ``include/symla_hip.h``.
* ``oracle/build/libsymla_oracle.so`` -- the CPU oracle (test infrastructure),
  built with g++ and OpenMP.

Usage: ``python -m ma57_amd.build`` (or ``python ma57_amd/build.py``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "ma57_amd" / "csrc"
PRODUCT = ROOT / "ma57_amd" / "libsymla_hip.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE = ORACLE_DIR / "build" / "libsymla_oracle.so"

HIP_SOURCES = ["factor_kernels.hip", "solve_kernels.hip", "cb_syrk.hip", "plan.cpp", "capi.cpp", "symbolic.cpp",
               "noalloc.hip", "mmio.cpp"]
LINK_LIBS: list[str] = []
# Packed FP32 ops are disabled for every HIP translation unit.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def build_product(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    objdir = ROOT / "ma57_amd" / "_objs"
    objdir.mkdir(exist_ok=True)
    cmds, objs = [], []
    for src in HIP_SOURCES:
        obj = objdir / (src + ".o")
        cmd = [hipcc, *HIP_FLAGS, "-I", str(ROOT / "include"), "-c", str(CSRC / src), "-o", str(obj)]
        if src.endswith(".cpp"):
            cmd.insert(1, "-x")
            cmd.insert(2, "hip")
        if verbose:
            print(" ".join(cmd))
        cmds.append(cmd)
        objs.append(str(obj))
    # translation units are independent: compile them side by side
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1, 8)) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = PRODUCT.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs, *LINK_LIBS])
    os.replace(tmp, PRODUCT)
    return PRODUCT


def build_oracle(verbose: bool = False) -> Path:
    (ORACLE_DIR / "build").mkdir(exist_ok=True)
    tmp = ORACLE.with_suffix(".so.tmp")
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-march=x86-64-v3",
           "-ffp-contract=off", "-o", str(tmp), str(ORACLE_DIR / "symla_oracle.cpp")]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, ORACLE)
    return ORACLE


FACADE_TEST = ROOT / "tests" / "cpp" / "facade_test"


def build_facade_test(verbose: bool = False) -> Path:
    """C++ program exercising include/symla/solver.hpp (run by the GPU tests)."""
    cmd = ["g++", "-O2", "-std=c++17", "-I", str(ROOT / "include"), str(ROOT / "tests" / "cpp" / "facade_test.cpp"),
           "-L", str(PRODUCT.parent), "-lsymla_hip", f"-Wl,-rpath,{PRODUCT.parent}", "-Wl,-rpath,$ORIGIN/../../ma57_amd",
           "-o", str(FACADE_TEST)]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    return FACADE_TEST


def build_all(verbose: bool = False) -> None:
    build_oracle(verbose)
    build_product(verbose)
    build_facade_test(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print(f"built {PRODUCT} and {ORACLE}")
