"""Builds the HIP engine library for gfx950 in-tree.

    python lennard-jonesium_amd/build_engine.py

Produces lennard-jonesium_amd/build/liblj_engine.so from csrc/lj_engine.hip.
Device code is compiled without implicit floating-point contraction (the
kernels write their fused multiply-adds explicitly) and without packed FP32
instructions.
"""

import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
OUT = HERE / "build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LJ_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-std=c++17", "-O3", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    f"-I{ROOT / 'include'}", f"-I{HERE / 'csrc'}",
]


HOST_SOURCES = ["initial_condition.cpp", "format.cpp", "simulation.cpp"]


def _run(cmd, verbose, what):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build of {what} failed")


def build(verbose: bool = False) -> pathlib.Path:
    """Builds liblj_engine.so (HIP kernels + C ABI) and liblj_host.so (host layer)."""
    OUT.mkdir(exist_ok=True)
    target = OUT / "liblj_engine.so"
    tmp = OUT / "liblj_engine.so.tmp"
    _run([HIPCC, *COMMON, str(HERE / "csrc" / "lj_engine.hip"), "-o", str(tmp)], verbose,
         "liblj_engine.so")
    os.replace(tmp, target)
    host = OUT / "liblj_host.so"
    htmp = OUT / "liblj_host.so.tmp"
    srcs = [str(HERE / "csrc" / "host" / f) for f in HOST_SOURCES]
    # the host layer sees only the plain-C headers: no HIP include path
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", f"-I{ROOT / 'include'}",
          *srcs, f"-L{OUT}", "-llj_engine",
          "-Wl,-rpath,$ORIGIN", "-o", str(htmp)], verbose, "liblj_host.so")
    os.replace(htmp, host)
    return target


if __name__ == "__main__":
    print(build(verbose=True))
