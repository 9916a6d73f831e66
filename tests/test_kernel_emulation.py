"""Host emulation of device-side index logic (no GPU): the register-blocked
NTT passes of kspir_device.h run thread by thread on the CPU and must equal
the oracle's NTT for each prime."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
N = 4096


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emul") / "ntt_emul.so")
    src = os.path.join(HERE, "native", "ntt_emulation.cpp")
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    subprocess.check_call([hipcc, "-O2", "-fPIC", "-shared", "-std=c++17", "-Xclang", "-target-feature",
                           "-Xclang", "-packed-fp32-ops", src, "-o", out], stderr=subprocess.DEVNULL)
    return ctypes.CDLL(out)


@pytest.mark.parametrize("q,root,modsel", [(268369921, 62736, 3), (249561089, 32309, 4), (16760833, 3689, 1)])
def test_register_blocked_ntt_matches_oracle(emul, oracle, q, root, modsel):
    rng = np.random.default_rng(q)
    for trial in range(4):
        a = rng.integers(0, q, N, dtype=np.uint32)
        if trial == 1:
            a[:] = q - 1
        if trial == 2:
            a[::2] = q - 1
            a[1::2] = 0
        x = a.copy()
        rc = emul.emul_ntt(x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.c_uint32(q),
                           ctypes.c_uint32(root), 0)
        assert rc == 0, "a lazy word left its bound between passes"
        want = oracle.ntt_forward(modsel, a.astype(np.uint64))
        assert np.array_equal(x.astype(np.uint64), want)
        rc = emul.emul_ntt(x.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.c_uint32(q),
                           ctypes.c_uint32(root), 1)
        assert rc == 0, "a lazy word left its bound between passes"
        assert np.array_equal(x, a)
