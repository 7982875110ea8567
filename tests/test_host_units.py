"""The library's host-only pieces under AddressSanitizer and UBSan (CPU):
tests/cpp/test_host_units.cpp (the serving arena's first-fit free list, the
autotuner's table parser) is built with g++ -fsanitize=address,undefined
and must run clean; and the C-ABI CPU suite passes against the library built
with its host code under both sanitizers."""
import os
import subprocess
import sys

import helpers

CSRC = os.path.join(helpers.ROOT, "openmp-blas_amd", "csrc")
SRC = os.path.join(helpers.ROOT, "tests", "cpp", "test_host_units.cpp")


def test_host_units_under_sanitizers(tmp_path):
    exe = str(tmp_path / "test_host_units")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-Werror",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-I" + CSRC, SRC, "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    out = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok "), out.stdout
    assert int(out.stdout.split()[1]) > 1000


def test_abi_suite_with_sanitized_host_code():
    """The C-ABI CPU suite (tests/test_abi_cpu.py: argument validation of
    every entry point, plan descriptor checks, serving-trace checks, the
    request sizing and the autotune-table loader) against the library built
    with its host code under AddressSanitizer and UBSan, in a process with
    the ASan runtime preloaded; any report fails the run."""
    pkg = os.path.join(helpers.ROOT, "openmp-blas_amd")
    subprocess.run(["make", "-j8", "asan"], cwd=pkg, check=True, capture_output=True)
    rt = subprocess.run(["make", "-s", "asan-runtime"], cwd=pkg, check=True, capture_output=True,
                        text=True).stdout.strip()
    assert os.path.exists(rt), rt
    preload = " ".join(p for p in (rt, os.environ.get("LD_PRELOAD", "")) if p)
    env = dict(os.environ, OMB_LIBRARY=os.path.join(pkg, "lib_asan", "libompblas_asan.so"),
               LD_PRELOAD=preload, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    out = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                          os.path.join(helpers.ROOT, "tests", "test_abi_cpu.py")],
                         capture_output=True, text=True, env=env, timeout=600)
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-3000:]
    assert "passed" in out.stdout and "failed" not in out.stdout
