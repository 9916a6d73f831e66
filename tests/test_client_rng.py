"""The rand() stream behind GPU client generation: the restatement of glibc's
generator and its jump-ahead (kspir_amd/csrc/kspir_rng.h) against the C
library's own rand(), and the per-client draw count of the library against
what the oracle's client generation actually consumes."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def rng_test(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("rng") / "rng_test")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "kspir_amd", "csrc"),
                           os.path.join(ROOT, "tests", "native", "rng_test.cpp"), "-o", exe])
    return exe


def test_rng_restatement_and_jumps_match_libc(rng_test):
    out = subprocess.run([rng_test], capture_output=True, text=True, check=True).stdout
    assert out.strip() == "ok"


@pytest.mark.parametrize("r,n1,seed", [(4, 64, 21), (16, 128, 5), (1, 256, 9)])
def test_client_draw_count_matches_oracle(oracle, rng_test, r, n1, seed):
    """after the oracle's client generation (srand(seed), secret, query, RGSW,
    BSGS and packing keys), libc's next rand() is output D of the stream,
    D = kspir_client_draws(r, n1)"""
    import kspir_amd
    d = ctypes.c_uint64(0)
    assert kspir_amd.lib.kspir_client_draws(r, n1, ctypes.byref(d)) == 0
    oracle.gen_client(r, n1, seed, 3, 0)
    libc = ctypes.CDLL(None)
    nxt = libc.rand()
    want = int(subprocess.run([rng_test, str(seed), str(d.value)], capture_output=True, text=True,
                              check=True).stdout)
    assert nxt == want
