"""The native runtime's host code under AddressSanitizer, LeakSanitizer and
UBSan, without a GPU.

hops_runtime.cpp is compiled with g++ for the host alone (the HIP API comes
from libamdhip64, the engine's device entry points from tests/asan/engine_stubs.c)
and driven by tests/asan/runtime_host_driver.cpp through the host-only hooks
of include/hops_runtime.h: systems, basis tables, the hierarchy and state
selection with every flux filter, the threshold helper, and the release of
every handle.  Any sanitizer report fails the test.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = "/opt/rocm"


def _need_toolchain():
    if not (shutil.which("g++") and shutil.which("gcc")):
        pytest.skip("g++ / gcc not found")
    if not os.path.exists(os.path.join(ROCM, "include", "hip", "hip_runtime.h")):
        pytest.skip("HIP headers not found")
    if not os.path.exists(os.path.join(ROCM, "lib", "libamdhip64.so")):
        pytest.skip("libamdhip64 not found")


def test_runtime_host_code_is_sanitizer_clean(tmp_path):
    _need_toolchain()
    san = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-O1"]
    inc = ["-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include")]
    rt = str(tmp_path / "runtime.o")
    stubs = str(tmp_path / "stubs.o")
    exe = str(tmp_path / "driver")
    steps = [
        ["g++", "-std=c++17", *san, *inc, "-c", os.path.join(ROOT, "mesohops_amd", "csrc", "hops_runtime.cpp"), "-o", rt],
        ["gcc", *san, "-c", os.path.join(ROOT, "tests", "asan", "engine_stubs.c"), "-o", stubs],
        ["g++", "-std=c++17", *san, *inc, os.path.join(ROOT, "tests", "asan", "runtime_host_driver.cpp"), rt, stubs,
         "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", exe],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.startswith("OK ")
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr
