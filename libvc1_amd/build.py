"""Build the in-tree shared libraries.

    python -m libvc1_amd.build           # HIP backend -> libvc1_amd/libvc1me.so
    python -m libvc1_amd.build --oracle  # also the CPU checkers under oracle/

The backend is compiled for gfx950 only with hipcc directly (no CMake); the
object code travels with the tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvc1me.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-mllvm", "-disable-machine-licm",
    # packed-FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{os.path.join(ROOT, 'include')}",
]

SOURCES = ["vc1me_kernels.hip", "vc1me_nn.hip", "vc1me_intensity.hip", "vc1me_api.cpp"]


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_backend(verbose: bool = True) -> str:
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"]
        _run([HIPCC, *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


ADAPTER_SRC = os.path.join(ROOT, "tests", "adapter", "adapter_test.cpp")
ADAPTER_BIN = os.path.join(ROOT, "tests", "adapter", "adapter_test")


def build_adapter_test() -> str:
    """The C++ test program of the libvc1 adapter (integration/hip_compute.h),
    compiled by the host compiler against include/vc1me.h and linked to the
    built library (tests/adapter/adapter_test)."""
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{os.path.join(ROOT, 'include')}",
          f"-I{os.path.join(HERE, 'integration')}", f"-I{os.path.dirname(ADAPTER_SRC)}", ADAPTER_SRC,
          f"-L{HERE}", "-l:libvc1me.so", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../libvc1_amd",
          "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-pthread", "-o", ADAPTER_BIN])
    return ADAPTER_BIN


def build_oracle(with_reference: bool = True) -> None:
    make = shutil.which("make") or "make"
    _run([make, "-s", "-C", os.path.join(ROOT, "oracle")])
    if with_reference and os.path.isdir("/root/reference"):
        _run([make, "-s", "-j8", "-C", os.path.join(ROOT, "oracle"), "ref"])


def main(argv: list[str]) -> int:
    build_backend()
    build_adapter_test()
    if "--oracle" in argv:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
