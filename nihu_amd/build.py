"""Build libnbem.so (HIP, gfx950) in-tree.

    python -m nihu_amd.build        # or __graft_entry__.build()

The host geometry and layout files are compiled with g++ -ffp-contract=off
(pair classification must round exactly as written); the HIP translation unit with
hipcc for gfx950 only.  Packed FP32 instructions are disabled for all device
code (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libnbem.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else "hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    host_o = []
    for name in ("nbem_geom", "nbem_layout"):
        o = os.path.join(BUILD, name + ".o")
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
              "-c", os.path.join(CSRC, name + ".cpp"), "-o", o])
        host_o.append(o)
    api_o = os.path.join(BUILD, "nbem_api.o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "nbem_api.hip"), "-o", api_o])
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", api_o, *host_o, "-o", tmp])
    os.replace(tmp, LIB)
    build_cpp_tests()
    return LIB


def build_cpp_tests():
    """The C++ front-end test program (tests/cpp), linked against libnbem.so."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "test_facade.cpp")
    if not os.path.exists(src):
        return
    out_dir = os.path.join(root, "tests", "cpp", "_build")
    os.makedirs(out_dir, exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "include"), src, "-o",
          os.path.join(out_dir, "test_facade"), "-L", HERE, "-lnbem", "-Wl,-rpath," + HERE,
          "-Wl,-rpath,$ORIGIN/../../../nihu_amd"])


if __name__ == "__main__":
    build()
    sys.exit(0)
