"""The host code under GCC's sanitizers (CPU only): the engine's geometry,
quadrature and Galerkin-layout routines (nbem_geom.cpp, nbem_layout.cpp)
run on one thread and then on four threads at once, with
-fsanitize=address,undefined (together with the FP64 oracle on four OpenMP
threads) and with -fsanitize=thread (host routines only: the oracle's
OpenMP critical sections are libgomp locks that the thread sanitizer does not
see, so it would report them as races).  The driver is
tests/cpp/sanitize_host.cpp; any sanitizer report or differing output fails."""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "cpp", "sanitize_host.cpp"),
       os.path.join(ROOT, "nihu_amd", "csrc", "nbem_geom.cpp"),
       os.path.join(ROOT, "nihu_amd", "csrc", "nbem_layout.cpp")]
ORACLE = os.path.join(ROOT, "oracle", "nbem_oracle.cpp")


def _build_and_run(tmp_path, flags, extra_src=(), defines=()):
    exe = str(tmp_path / "sanitize_host")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer", "-ffp-contract=off", *flags,
           *["-D" + d for d in defines], *SRC, *extra_src, "-o", exe, "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0 and "OK" in r.stdout, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "WARNING: ThreadSanitizer" not in r.stderr


def test_address_and_undefined_behaviour(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=address,undefined", "-fopenmp"], [ORACLE], ["WITH_ORACLE"])


def test_threads(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=thread"])
