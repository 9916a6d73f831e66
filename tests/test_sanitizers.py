"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU):
the CPU oracle, the rand() restatement (kspir_rng.h) and the host-only paths
of the C++ drop-in (kspir_pir.cpp) are built with -fsanitize=address,undefined
and run on small inputs (tests/native/sanitize_host.cpp); a report fails the
program."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_host_code_under_asan_ubsan(tmp_path):
    lib = os.path.join(ROOT, "kspir_amd", "libkspir_amd.so")
    if not os.path.exists(lib):
        pytest.fail("kspir_amd/libkspir_amd.so not built (python kspir_amd/build.py)")
    exe = str(tmp_path / "sanitize_host")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-fopenmp", os.path.join(HERE, "native", "sanitize_host.cpp"),
           os.path.join(ROOT, "oracle", "kspir_oracle.cpp"), os.path.join(ROOT, "kspir_amd", "cpp", "kspir_pir.cpp"),
           "-I", os.path.join(ROOT, "kspir_amd", "cpp"), "-I", os.path.join(ROOT, "kspir_amd", "csrc"),
           "-I", os.path.join(ROOT, "oracle"), "-L", os.path.join(ROOT, "kspir_amd"), "-lkspir_amd",
           "-Wl,-rpath," + os.path.join(ROOT, "kspir_amd"), "-o", exe]
    subprocess.check_call(cmd)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert "sanitize ok" in out.stdout
