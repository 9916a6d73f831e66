"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
import this module.  It loads oracle/liboracle.so (built by oracle/Makefile)
and, when present, oracle/_ref/libref_crt.so (the reference's own crt.cpp).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_crt.so")

N = 4096
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def build(ref=False):
    """Compile the oracle (and the reference crt.cpp when ref=True)."""
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.check_call(["make", "-s", "-C", HERE] + targets)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.ko_const.restype = ctypes.c_uint64
        L.ko_qinv.restype = ctypes.c_uint64
        L.ko_min_root.restype = ctypes.c_uint64
        L.ko_min_root.argtypes = [ctypes.c_uint64]
        L.ko_crt_compose.restype = ctypes.c_uint64
        L.ko_crt_compose.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own crt.cpp as a shared library, or None if not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        R = ctypes.CDLL(REF_LIB)
        R.ref_crt_compose.restype = ctypes.c_uint64
        R.ref_crt_compose.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        _ref = R
    return _ref


def p64(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


def p32(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i32p)


def const(i):
    return int(lib().ko_const(i))


Q1, Q2, Q, BS, P, DELTA = (const(i) for i in (1, 2, 3, 4, 5, 6))


def set_threads(n):
    lib().ko_set_threads(int(n))


def shapes(r, n1):
    n2 = N // 2 // n1
    lr = max(0, r.bit_length() - 1)
    return {
        "secret": (N,),
        "db": (r, N * N // 2),
        "query": (2, 2, N),
        "rgsw": (4, 2, N),
        "baby": (n1 // 2, 6, 2, N),
        "giant": (max(n2 - 1, 0), 2, 2, N),
        "packing": (lr, 4, 2, N),
        "perms": (n1, N),
    }


class Instance:
    """Inputs of one online answer, generated in test-pir order."""

    def __init__(self, r, n1, **arrays):
        self.r, self.n1 = r, n1
        self.n2 = N // 2 // n1
        for k, v in arrays.items():
            setattr(self, k, v)


def gen_instance(r=16, n1=128, seed=1, plant=True, db=None):
    """test-pir.cpp:12-98 with srand(seed) (seed 1 == glibc default stream)."""
    sh = shapes(r, n1)
    a = {k: np.zeros(v, dtype=np.int32 if k == "perms" else np.uint64) for k, v in sh.items() if k != "db"}
    a["db"] = db if db is not None else np.empty(sh["db"], dtype=np.uint64)
    target = np.zeros(2, dtype=np.int32)
    rc = lib().ko_gen_instance(r, n1, ctypes.c_uint(seed), int(plant), p64(a["secret"]), p32(target),
                               p64(a["db"]), p64(a["query"]), p64(a["rgsw"]), p64(a["baby"]),
                               p64(a["giant"]), p64(a["packing"]), p32(a["perms"]))
    assert rc == 0
    return Instance(r, n1, target=(int(target[0]), int(target[1])), **a)


def gen_client(r, n1, seed, tcol, tpack):
    sh = shapes(r, n1)
    a = {k: np.zeros(sh[k], dtype=np.uint64) for k in ("secret", "query", "rgsw", "baby", "giant", "packing")}
    rc = lib().ko_gen_client(r, n1, ctypes.c_uint(seed), tcol, tpack, p64(a["secret"]), p64(a["query"]),
                             p64(a["rgsw"]), p64(a["baby"]), p64(a["giant"]), p64(a["packing"]))
    assert rc == 0
    return a


def gen_db(r, n1, seed, plant=False, tcol=0, tpack=0, out=None):
    db = out if out is not None else np.empty((r, N * N // 2), dtype=np.uint64)
    rc = lib().ko_gen_db(r, n1, ctypes.c_uint(seed), int(plant), tcol, tpack, p64(db))
    assert rc == 0
    return db


def perm_matrix(n1):
    out = np.zeros((n1, N), dtype=np.int32)
    lib().ko_permutation_matrix(p32(out), n1)
    return out


def answer(inst=None, *, r=None, n1=None, db=None, query=None, rgsw=None, baby=None, giant=None,
           packing=None, perms=None, want_stages=False):
    """Reference online answer; returns response [2][N] (and stage arrays)."""
    if inst is not None:
        r, n1 = inst.r, inst.n1
        db, query, rgsw, baby, giant, packing, perms = (inst.db, inst.query, inst.rgsw, inst.baby,
                                                        inst.giant, inst.packing, inst.perms)
    resp = np.zeros((2, N), dtype=np.uint64)
    cbuf = fd = ext = None
    null = ctypes.cast(None, _u64p)
    if want_stages:
        cbuf = np.zeros((N, n1, 2), dtype=np.uint64)
        fd = np.zeros((r, 2, N), dtype=np.uint64)
        ext = np.zeros((r, 2, N), dtype=np.uint64)
    rc = lib().ko_answer(r, n1, p64(db), p64(query), p64(rgsw), p64(baby), p64(giant), p64(packing),
                         p32(perms), p64(resp), p64(cbuf) if want_stages else null,
                         p64(fd) if want_stages else null, p64(ext) if want_stages else null)
    assert rc == 0
    if want_stages:
        return resp, {"cipherbuf": cbuf, "firstdim": fd, "ext": ext}
    return resp


def gen_records(r, n1, seed=1, plant=False, tcol=0, tpack=0):
    """records [r][N][N/2] (database_tobsgsntt input) and the words [r][N*N/2]
    of the same database, generated as gen_db does"""
    rec = np.zeros((r, N, N // 2), dtype=np.uint64)
    db = np.zeros((r, N * N // 2), dtype=np.uint64)
    rc = lib().ko_gen_records(r, n1, seed, int(plant), tcol, tpack, p64(rec), p64(db))
    assert rc == 0
    return rec, db


def inverse_encode(col):
    """inverse_encode of one record column (N values < p)"""
    out = np.ascontiguousarray(col, dtype=np.uint64).copy()
    lib().ko_inverse_encode(p64(out))
    return out


def records_to_words(r, n1, records, block_stride=None):
    """words [r][N*N/2] of records [r or 1][N][N/2] (block_stride 0 repeats block 0)"""
    rec = np.ascontiguousarray(records, dtype=np.uint64)
    stride = N * N // 2 if block_stride is None else block_stride
    out = np.zeros((r, N * N // 2), dtype=np.uint64)
    rc = lib().ko_records_to_words(r, n1, p64(rec), ctypes.c_int64(stride), p64(out))
    assert rc == 0
    return out


def database_tobsgsntt(n1, data):
    """database_tobsgsntt of one block: data [N][N/2] -> [N/2][N] mod Q"""
    out = np.zeros((N // 2, N), dtype=np.uint64)
    rc = lib().ko_database_tobsgsntt(n1, p64(np.ascontiguousarray(data, dtype=np.uint64)), p64(out))
    assert rc == 0
    return out


def answer_parallel(inst, count, threads):
    """`count` independent answers of inst's query on `threads` threads
    (the all-cores CPU baseline); returns the response of the first"""
    resp = np.zeros((2, N), dtype=np.uint64)
    rc = lib().ko_answer_parallel(count, threads, inst.r, inst.n1, p64(inst.db), p64(inst.query), p64(inst.rgsw),
                                  p64(inst.baby), p64(inst.giant), p64(inst.packing), p32(inst.perms), p64(resp))
    assert rc == 0, rc
    return resp


def last_stage_us():
    out = (ctypes.c_double * 5)()
    lib().ko_last_stage_us(out)
    return list(out)


def decrypt(r, secret, response):
    msg = np.zeros(N, dtype=np.uint64)
    resp = np.ascontiguousarray(response, dtype=np.uint64)
    lib().ko_decrypt(r, p64(np.ascontiguousarray(secret)), p64(resp), p64(msg))
    return msg


def ntt_forward(modsel, a):
    x = np.ascontiguousarray(a, dtype=np.uint64).copy()
    lib().ko_ntt_forward(modsel, p64(x))
    return x


def ntt_inverse(modsel, a):
    x = np.ascontiguousarray(a, dtype=np.uint64).copy()
    lib().ko_ntt_inverse(modsel, p64(x))
    return x


def decompose(x, ell, base, bg, q):
    out = np.zeros((ell, N), dtype=np.uint64)
    lib().ko_decompose(p64(out), p64(np.ascontiguousarray(x, dtype=np.uint64)), ell,
                       ctypes.c_uint64(base), ctypes.c_uint64(bg), ctypes.c_uint64(q))
    return out


def decompose_rlwe(a, b, ell, base, bg, q):
    out = np.zeros((2 * ell, N), dtype=np.uint64)
    lib().ko_decompose_rlwe(p64(out), p64(np.ascontiguousarray(a, dtype=np.uint64)),
                            p64(np.ascontiguousarray(b, dtype=np.uint64)), ell,
                            ctypes.c_uint64(base), ctypes.c_uint64(bg), ctypes.c_uint64(q))
    return out


def decompose_bsgs(xq, xbs, ell=3, base=1 << 20, bg=1 << 20):
    oq = np.zeros((ell, N), dtype=np.uint64)
    ob = np.zeros((ell, N), dtype=np.uint64)
    lib().ko_decompose_bsgs(p64(oq), p64(ob), p64(np.ascontiguousarray(xq, dtype=np.uint64)),
                            p64(np.ascontiguousarray(xbs, dtype=np.uint64)), ell,
                            ctypes.c_uint64(base), ctypes.c_uint64(bg))
    return oq, ob


def crt_inv_bsgs(xq, xbs):
    out = np.zeros((N, 2), dtype=np.uint64)
    lib().ko_crt_inv_bsgs(p64(out), p64(np.ascontiguousarray(xq, dtype=np.uint64)),
                          p64(np.ascontiguousarray(xbs, dtype=np.uint64)))
    return [int(lo) | (int(hi) << 64) for lo, hi in out]


def hexl_ntt_index():
    out = np.zeros(N, dtype=np.int32)
    lib().ko_hexl_ntt_index(p32(out))
    return out


def query_encode():
    out = np.zeros(N, dtype=np.int32)
    lib().ko_query_encode(p32(out))
    return out


def automorphic_rns(x, index, q):
    out = np.zeros(N, dtype=np.uint64)
    lib().ko_automorphic_rns(p64(out), p64(np.ascontiguousarray(x, dtype=np.uint64)), index,
                             ctypes.c_uint64(q))
    return out


def db_product_scalar(db_block, cbuf, n1, n2):
    out = np.zeros((n2, 2 * N), dtype=np.uint64)
    lib().ko_db_product_scalar(p64(out), p64(np.ascontiguousarray(db_block)), p64(np.ascontiguousarray(cbuf)), n1, n2)
    return out


def db_product_vectorised():
    """whether db_product takes the AVX-512 loop on this host"""
    return bool(lib().ko_db_product_vectorised())


def ntt_set_vector(on):
    """NTTs 8 lanes at a time with AVX-512 (on) or scalar; True when this
    host has AVX-512F/DQ (without it the scalar form runs either way)"""
    return bool(lib().ko_ntt_set_vector(1 if on else 0))


def db_product(db_block, cbuf, n1, n2):
    out = np.zeros((n2, 2 * N), dtype=np.uint64)
    lib().ko_db_product(p64(out), p64(np.ascontiguousarray(db_block)), p64(np.ascontiguousarray(cbuf)), n1, n2)
    return out


def ref_database_tocrt(data_ntt, n1):
    """the reference's own database_tocrt (crt.cpp:859-885) of one block:
    data_ntt [N/2][N] -> datacrt words [N*N/2]"""
    R = ref_lib()
    out = np.zeros(N * N // 2, dtype=np.uint64)
    R.ref_database_tocrt(p64(out), p64(np.ascontiguousarray(data_ntt, dtype=np.uint64)), n1)
    return out


def ref_db_product(db_block, cbuf, n1, n2):
    R = ref_lib()
    out = np.zeros((n2, 2 * N), dtype=np.uint64)
    R.ref_db_product(p64(out), p64(np.ascontiguousarray(db_block)), p64(np.ascontiguousarray(cbuf)), n1, n2)
    return out


def crt_inv_bsgs_aux(xq, xaux):
    out = np.zeros((N, 2), dtype=np.uint64)
    lib().ko_crt_inv_bsgs_aux(p64(out), p64(np.ascontiguousarray(xq, dtype=np.uint64)),
                              p64(np.ascontiguousarray(xaux, dtype=np.uint64)))
    return [int(lo) | (int(hi) << 64) for lo, hi in out]


def decompose_bsgs_aux(xq, xaux, ell=5, base=1 << 14, bg=1 << 14):
    oq = np.zeros((ell, N), dtype=np.uint64)
    oa = np.zeros((ell, N), dtype=np.uint64)
    lib().ko_decompose_bsgs_aux(p64(oq), p64(oa), p64(np.ascontiguousarray(xq, dtype=np.uint64)),
                                p64(np.ascontiguousarray(xaux, dtype=np.uint64)), ell,
                                ctypes.c_uint64(base), ctypes.c_uint64(bg))
    return oq, oa


def gen_offline_giant(n1, seed):
    """srand(seed), Secret(crtMod), then the (crtMod, auxMod) offline key
    (ell 5, 2^14, 2^14) and auto key (ell 3, 2^11*auxMod, 2^15) at 5^N1"""
    secret = np.zeros(N, dtype=np.uint64)
    off = np.zeros((10, 2, N), dtype=np.uint64)
    auto = np.zeros((6, 2, N), dtype=np.uint64)
    rc = lib().ko_gen_offline_giant(n1, ctypes.c_uint(seed), p64(secret), p64(off), p64(auto))
    assert rc == 0
    return secret, off, auto


def giant_from_offline(n1, offline2, auto2, want_chain=True):
    n2 = N // 2 // n1
    chain = np.zeros((max(n2 - 2, 0), 6, 2, N), dtype=np.uint64) if want_chain else None
    giant = np.zeros((n2 - 1, 3, 2, N), dtype=np.uint64)
    null = ctypes.cast(None, _u64p)
    rc = lib().ko_giant_from_offline(n1, p64(np.ascontiguousarray(offline2, dtype=np.uint64)),
                                     p64(np.ascontiguousarray(auto2, dtype=np.uint64)),
                                     p64(chain) if want_chain and chain.size else null, p64(giant))
    assert rc == 0
    return chain, giant


def rlwe_phase(modsel, secret, a, b, ntt_form):
    out = np.zeros(N, dtype=np.uint64)
    rc = lib().ko_rlwe_phase(modsel, p64(np.ascontiguousarray(secret, dtype=np.uint64)),
                             p64(np.ascontiguousarray(a, dtype=np.uint64)),
                             p64(np.ascontiguousarray(b, dtype=np.uint64)), int(ntt_form), p64(out))
    assert rc == 0
    return out
