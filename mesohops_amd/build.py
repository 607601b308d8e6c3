"""Build libhops_mi355x.so in-tree for gfx950 with hipcc.

Usage: python -m mesohops_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "csrc", f) for f in ("hops_engine.hip", "noise_fft.hip", "hops_select.hip", "hops_runtime.cpp")]
OUT = os.path.join(HERE, "libhops_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
         # packed FP32 instructions are kept out of every kernel
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(HERE, "..", "include")]


def build(verbose=False):
    """Compile each source to an object in parallel (one hipcc per file), then
    link the shared library."""
    tmp = OUT + ".tmp"
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in FLAGS if f != "-shared"]
    procs = []
    objs = []
    for src in SRC:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        procs.append((src, subprocess.Popen([HIPCC] + cflags + ["-c", src, "-o", obj], stdout=subprocess.PIPE,
                                            stderr=subprocess.PIPE, text=True)))
    failed = []
    for src, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(out + err)
        elif verbose:
            sys.stderr.write(err)
    if failed:
        raise RuntimeError("hipcc failed on " + ", ".join(os.path.basename(f) for f in failed))
    r = subprocess.run([HIPCC] + FLAGS + objs + ["-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed (%d)" % r.returncode)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
