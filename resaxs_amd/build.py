"""Build recipe for libresaxs_amd.so (gfx950), in-tree next to this file.

hipcc compiles the kernels and the host engine; packed-FP32 instructions are
disabled on the device side (they are not used by any kernel here).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libresaxs_amd.so")
ARCH = os.environ.get("RSX_OFFLOAD_ARCH", "gfx950")

SOURCES = ["kernels.hip", "engine.cpp", "pdb.cpp", "fit.cpp", "noalloc.hip"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    """Compile every source and link the shared library; always rebuilds."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    flags = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function",
             "-Wno-unused-command-line-argument"]
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + flags + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            raise RuntimeError("compile failed: " + " ".join(cmd))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


COMPAT_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "compat_test.cpp")
COMPAT_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "compat_test")
FIT_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "fit_test.cpp")
FIT_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "fit_test")


def _build_cpp_program(src, binary, verbose):
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-std=c++14", "-O2", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I" + os.path.join(ROOT, "include"),
           src, "-o", binary + ".tmp", "-L" + HERE, "-lresaxs_amd",
           "-Wl,-rpath,$ORIGIN/../../resaxs_amd", "-Wl,--allow-shlib-undefined"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(binary + ".tmp", binary)
    return binary


def build_compat_test(verbose=False):
    """Compile the C++ programs that drive include/resaxs/ (the resaxslib-compatible
    surface and the calc_saxs fitting driver) against libresaxs_amd.so, as C++14 host
    programs (the reference's language level); g++ only, no device code."""
    _build_cpp_program(FIT_TEST_SRC, FIT_TEST_BIN, verbose)
    return _build_cpp_program(COMPAT_TEST_SRC, COMPAT_TEST_BIN, verbose)


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_compat_test(verbose=True))
