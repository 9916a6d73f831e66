"""Build libkspir_amd.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "libkspir_amd.so")
SOURCES = [os.path.join(HERE, "csrc", "kspir_capi.hip"), os.path.join(HERE, "csrc", "kspir_tune.cpp"),
           os.path.join(HERE, "cpp", "kspir_pir.cpp")]
DEPS = SOURCES + [os.path.join(HERE, "csrc", f) for f in ("kspir_device.h", "kspir_kernels.h")] + [os.path.join(HERE, "cpp", "kspir_pir.h")] + [
    os.path.join(ROOT, "include", f) for f in ("kspir_amd.h", "kspir_amd_runtime.h")]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_dropin_test(verbose=False):
    """tests/native/dropin_test: the test-pir online calls against kspir_pir.h"""
    outdir = os.path.join(ROOT, "tests", "native", "_build")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "dropin_test")
    cmd = ["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "native", "dropin_test.cpp"),
           "-I", os.path.join(HERE, "cpp"), "-L", HERE, "-lkspir_amd",
           "-Wl,-rpath," + os.path.relpath(HERE, outdir).join(["$ORIGIN/", ""]), "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


REF_TEST_PIR = "/root/reference/tests/test-pir.cpp"


def build_reference_driver(verbose=False):
    """tests/native/_build/test_pir: the reference's own tests/test-pir.cpp,
    compiled from where it lies, unchanged, against the drop-in headers
    (kspir_amd/cpp/compat: pir.h, crt.h) and linked with libkspir_amd.so.
    Skipped when the reference tree is absent; the binary is never committed."""
    if not os.path.exists(REF_TEST_PIR):
        return None
    outdir = os.path.join(ROOT, "tests", "native", "_build")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "test_pir")
    cmd = ["g++", "-O2", "-std=c++17", "-w", REF_TEST_PIR, "-I", os.path.join(HERE, "cpp", "compat"),
           "-I", os.path.join(HERE, "cpp"), "-L", HERE, "-lkspir_amd",
           "-Wl,-rpath," + os.path.relpath(HERE, outdir).join(["$ORIGIN/", ""]), "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


def build(verbose=False):
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-result", "-Wno-unused-value", "-I", os.path.join(ROOT, "include"),
           *SOURCES, "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    build_dropin_test(verbose)
    build_reference_driver(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    print("built", OUT)
