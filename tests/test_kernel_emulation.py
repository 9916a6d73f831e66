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


Q = 268369921 * 249561089
BS = 16760833


@pytest.fixture(scope="module")
def digits(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("digits") / "digit_emul.so")
    src = os.path.join(HERE, "native", "digit_emulation.cpp")
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    subprocess.check_call([hipcc, "-O2", "-fPIC", "-shared", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", src, "-o", out],
                          stderr=subprocess.DEVNULL)
    return ctypes.CDLL(out)


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _tfhe_offset(ell, basebits, gbits):
    """utils.cpp:754-811: the offset added before the d > q test"""
    off = sum(1 << (basebits + i * gbits) for i in range(ell)) << (gbits - 1)
    return off + ((1 << basebits) // 2 if basebits else 0)


def _inputs(edges, hi, seed):
    """N inputs: the boundary values first, uniform values in [0, hi) after"""
    rng = np.random.default_rng(seed)
    x = np.array([int(v) for v in rng.integers(0, hi, N, dtype=np.uint64)], dtype=object)
    for i, e in enumerate(edges):
        x[i] = e
    return x


@pytest.mark.parametrize("which,ell,basebits,gbits", [("giant", 2, 20, 18), ("packing", 4, 0, 14)])
def test_device_tfhe_digits_at_the_boundaries(digits, oracle, which, ell, basebits, gbits):
    """tfhe_digit (giant steps and packing) equals the oracle's decompose on
    x = 0, x = Q (the q - 0 word an automorphism leaves), x + offset = Q
    (where d > q is false) and = Q + 1 (true), and random x"""
    off = _tfhe_offset(ell, basebits, gbits)
    edges = [0, 1, Q - 1, Q, Q - off, Q - off + 1, Q - off - 1, off, Q - off + (1 << basebits)]
    x = np.array(_inputs(edges, Q, ell), dtype=np.uint64)
    got = np.zeros((ell, N), dtype=np.uint64)
    getattr(digits, f"emul_digits_{which}")(_p(x), N, _p(got))
    want = oracle.decompose(x, ell, 1 << basebits if basebits else 0, 1 << gbits, Q)
    assert np.array_equal(got, want)
    # the two sides of the d > q test give different digits, so the test is live
    assert not np.array_equal(got[:, 4], got[:, 5])


def test_device_signed_digits_at_the_boundaries(digits, oracle):
    """signed_digit (external product) equals the oracle's decompose_rlwe at
    x = 0, Q/2, Q/2 + 1 (the sign flip), Q - 1 and at random x"""
    h = Q >> 1
    edges = [0, 1, h - 1, h, h + 1, Q - 1, (1 << 19) - 1, 1 << 19, (1 << 20) - 1, 1 << 20, Q - (1 << 20)]
    x = np.array(_inputs(edges, Q, 3), dtype=np.uint64)
    got = np.zeros((2, N), dtype=np.uint64)
    digits.emul_digits_ext(_p(x), N, _p(got))
    want = oracle.decompose_rlwe(x, x, 2, 1 << 20, 1 << 18, Q)
    assert np.array_equal(got, want[:2]) and np.array_equal(got, want[2:])


def test_device_bsgs_digits_at_the_boundaries(digits, oracle):
    """bsgs_digit (baby steps) of the composed value v equals the oracle's
    decompose_bsgs of (v mod Q, v mod bsMod) at v = 0, M - 1, v + offset = M
    (where d > M is false) and = M + 1 (true), and at random v"""
    M = Q * BS
    off = (((1 << 20) + (1 << 40) + (1 << 60)) << 19) | (1 << 19)
    edges = [0, 1, M - 1, M - off, M - off + 1, M - off - 1, off, Q, Q - 1, M - Q]
    v = _inputs(edges, 1 << 62, 5)
    rng = np.random.default_rng(9)
    for i in range(len(edges), N):   # spread the random values over all of [0, M)
        v[i] = (int(v[i]) * 1024 + int(rng.integers(0, 1024))) % M
    lo = np.array([int(t) & ((1 << 64) - 1) for t in v], dtype=np.uint64)
    hi = np.array([int(t) >> 64 for t in v], dtype=np.uint64)
    gq, gb = np.zeros((3, N), dtype=np.uint64), np.zeros((3, N), dtype=np.uint64)
    digits.emul_digits_bsgs(_p(lo), _p(hi), N, _p(gq), _p(gb))
    xq = np.array([int(t) % Q for t in v], dtype=np.uint64)
    xb = np.array([int(t) % BS for t in v], dtype=np.uint64)
    wq, wb = oracle.decompose_bsgs(xq, xb)
    assert np.array_equal(gq, wq) and np.array_equal(gb, wb)
    assert not (np.array_equal(gq[:, 3], gq[:, 4]) and np.array_equal(gb[:, 3], gb[:, 4]))
