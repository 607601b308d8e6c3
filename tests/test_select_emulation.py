"""The device form of the adaptive basis selection (hops_select.hip), run on
the CPU: the native runtime and the selection kernels are compiled with g++
against a small emulation of the HIP calls they use (tests/emu/: one
std::thread per work-item, barriers for __syncthreads, host memory for
device memory filled with a stale pattern, a stable sort standing in for
rocPRIM's segmented sorts) under AddressSanitizer and UBSan.  Two batches of
the same trajectories, one selecting on the (emulated) device and one on the
host threads, must keep bitwise-equal state lists, auxiliary lists and Phi
after every call, with no sanitizer report.  The engine's other device entry
points are CPU stand-ins with a toy step that grows the bases
(tests/emu/emu_engine.cpp); the selection arithmetic itself is the product
code."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("emu")
    san = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-O1", "-std=c++20",
           "-I", os.path.join(EMU, "include")]
    objs = []
    srcs = [(os.path.join(ROOT, "mesohops_amd", "csrc", "hops_select.hip"), ["-x", "c++"]),
            (os.path.join(ROOT, "mesohops_amd", "csrc", "hops_runtime.cpp"), []),
            (os.path.join(EMU, "emu_engine.cpp"), [])]
    procs = []
    for src, extra in srcs:
        obj = str(d / (os.path.basename(src) + ".o"))
        objs.append(obj)
        procs.append(subprocess.Popen(["g++", *san, *extra, "-c", src, "-o", obj], stdout=subprocess.PIPE,
                                      stderr=subprocess.PIPE, text=True))
    for p in procs:
        out, err = p.communicate()
        assert p.returncode == 0, err[-4000:]
    exe = str(d / "select_driver")
    r = subprocess.run(["g++", *san, os.path.join(EMU, "select_driver.cpp"), *objs, "-lpthread", "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


# n sites, steps per call, delta_a, delta_s, Markovian filter, slots, static basis
CASES = {
    "hierarchy_only": (10, 3, 5e-4, 0.0, 0, 3, 0),
    "states_and_hierarchy": (12, 3, 1e-3, 1e-3, 0, 3, 0),
    "markov": (8, 3, 1e-3, 5e-4, 1, 2, 0),
    # every slot starts from a STATIC_BASIS: the first step reads psi through
    # the device descriptors before any selection has uploaded them
    "static_basis": (8, 3, 1e-3, 1e-3, 0, 3, 1),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_selection_emulated_equals_host(driver, name):
    n, steps, da, ds, mk, B, st = CASES[name]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([driver, str(n), str(steps), str(da), str(ds), str(mk), str(B), str(st)], capture_output=True,
                       text=True, env=env, timeout=900)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert r.stdout.startswith("OK ")
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr
