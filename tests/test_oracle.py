"""Pins of the CPU oracle (oracle/kspir_oracle.cpp) against the reference's own
known answers and against what the mathematics fixes.  CPU only."""
import numpy as np
import pytest

N = 4096


def bitrev(x, bits=12):
    return int(f"{x:0{bits}b}"[::-1], 2)


# ---------------------------------------------------------------- constants

def test_parameter_constants(oracle):
    O = oracle
    q1, q2, Q, bs, p, delta = O.Q1, O.Q2, O.Q, O.BS, O.P, O.DELTA
    assert (q1, q2, bs, p) == (268369921, 249561089, 16760833, 65537)
    assert Q == 66974689739603969
    assert delta == round(Q / p) == (Q + p // 2) // p        # params.h:98 bsgsDelta
    bs_inv_q = 51309461009346113                              # crt.h:33 bsModInv
    assert bs_inv_q * bs % Q == 1
    assert 3920321 * Q % bs == 1                              # crt.h:34 crtModInv
    # crt.h:60-61: 2^128 mod (Q * bsMod) used by compute_mult_qinvqi_bsgs
    assert (54824 << 63) + 1764040975123512121 == pow(2, 128, Q * bs)
    # crt.cpp:25-28 Barrett / CRT constants
    assert 163640210 * q2 % q1 == 1 and 97389680 * q1 % q2 == 1
    assert (275 << 64) + 7906011006380390721 == (1 << 128) // Q
    assert 68736257792 == (1 << 64) // q1 and 73916747789 == (1 << 64) // q2   # crt.h:20-21


def test_qinv_table(oracle):
    # utils.cpp:172-197: QInv(r, crtMod) * r == 1 mod Q where the table has an entry
    for r in (1, 2, 4, 8, 16, 32, 64, 256, 512, 4096):
        v = int(oracle.lib().ko_qinv(r))
        assert v * r % oracle.Q == 1, r
    for r in (128, 1024, 2048):   # absent from the table: the reference returns 0 (cannot decrypt)
        assert int(oracle.lib().ko_qinv(r)) == 0


# --------------------------------------------------------------------- roots

def test_ntt_roots(oracle):
    O = oracle
    root_q = 3375402822066082                                  # crt.h:81
    assert pow(root_q, N, O.Q) == O.Q - 1
    # HEXL's composite-modulus NTT is the CRT of the minimal-root prime NTTs
    assert int(O.lib().ko_min_root(O.Q1)) == root_q % O.Q1 == 62736
    assert int(O.lib().ko_min_root(O.Q2)) == root_q % O.Q2 == 32309
    assert O.const(8) == 3689 and O.const(9) == 13             # bsMod, p minimal roots
    for q, w in ((O.BS, 3689), (O.P, 13)):
        assert pow(w, N, q) == q - 1
        assert all(pow(x, N, q) != q - 1 for x in range(2, w))


@pytest.mark.parametrize("vector", [True, False])
@pytest.mark.parametrize("modsel,q,root", [(0, 66974689739603969, 3375402822066082),
                                           (1, 16760833, 3689), (2, 65537, 13),
                                           (3, 268369921, 62736), (4, 249561089, 32309)])
def test_ntt_closed_form(oracle, modsel, q, root, vector):
    """slot i of ComputeForward holds a(root^(2*bitrev(i)+1)) (utils.cpp:1476-1490),
    for the AVX-512 lane form and the scalar form of the oracle's NTT."""
    if not oracle.ntt_set_vector(vector) and vector:
        pytest.skip("host without AVX-512: the scalar NTT is the only one")
    try:
        rng = np.random.default_rng(modsel)
        a = rng.integers(0, q, N, dtype=np.uint64)
        fa = oracle.ntt_forward(modsel, a)
        ai = [int(v) for v in a]
        for i in list(range(4)) + [1000, 2048, 4095]:
            x = pow(root, 2 * bitrev(i) + 1, q)
            acc, xp = 0, 1
            for c in ai:
                acc += c * xp
                xp = xp * x % q
            assert int(fa[i]) == acc % q, (modsel, i)
        assert np.array_equal(oracle.ntt_inverse(modsel, fa), a)
    finally:
        oracle.ntt_set_vector(True)


def test_ntt_lanes_match_scalar(oracle):
    """the oracle's AVX-512 NTT (the lane form its parity checks and the CPU
    baseline run) gives the scalar NTT's words, forward and inverse, for
    every modulus, on random inputs with 0, q - 1 and the non-canonical q
    (which enters as 0)"""
    if not oracle.ntt_set_vector(True):
        pytest.skip("host without AVX-512: the scalar NTT is the only one")
    qs = {0: oracle.Q, 1: oracle.BS, 2: oracle.P, 3: oracle.Q1, 4: oracle.Q2, 5: 268361729}   # 5: auxMod, crt.h:37
    rng = np.random.default_rng(11)
    try:
        for modsel in range(6):
            q = qs[modsel]
            for trial in range(6):
                a = rng.integers(0, q, N, dtype=np.uint64)
                a[[0, 1, 4095]] = [0, q - 1, q]
                a[rng.integers(0, N, 16)] = q
                for fn in (oracle.ntt_forward, oracle.ntt_inverse):
                    oracle.ntt_set_vector(True)
                    lanes = fn(modsel, a)
                    oracle.ntt_set_vector(False)
                    scalar = fn(modsel, a)
                    assert np.array_equal(lanes, scalar), (modsel, trial, fn.__name__)
    finally:
        oracle.ntt_set_vector(True)


def test_ntt_negacyclic_convolution(oracle):
    """INTT(NTT(a) * NTT(b)) equals the schoolbook product mod X^N + 1 (sparse b)."""
    Q = oracle.Q
    rng = np.random.default_rng(7)
    a = rng.integers(0, Q, N, dtype=np.uint64)
    b = np.zeros(N, dtype=np.uint64)
    terms = {0: 5, 17: Q - 3, 4095: 123456789}
    for k, v in terms.items():
        b[k] = v
    fa, fb = oracle.ntt_forward(0, a), oracle.ntt_forward(0, b)
    prod = np.array([int(x) * int(y) % Q for x, y in zip(fa, fb)], dtype=np.uint64)
    got = oracle.ntt_inverse(0, prod)
    want = [0] * N
    for k, v in terms.items():
        for i in range(N):
            j = i + k
            if j < N:
                want[j] = (want[j] + int(a[i]) * v) % Q
            else:
                want[j - N] = (want[j - N] - int(a[i]) * v) % Q
    assert [int(x) for x in got] == want


def test_hexl_ntt_index_closed_form(oracle):
    h = oracle.hexl_ntt_index()
    assert list(h[:4]) == [1, 4097, 2049, 6145]                    # utils.cpp:1478 comment
    assert all(int(h[i]) == 2 * bitrev(i) + 1 for i in range(N))


def test_permutation_realises_automorphism(oracle):
    """NTT(tau_{5^i}(a)) == NTT(a)[perm[i]] (test-utils.cpp:361-404)."""
    Q = oracle.Q
    perms = oracle.perm_matrix(128)
    rng = np.random.default_rng(3)
    a = rng.integers(0, Q, N, dtype=np.uint64)
    fa = oracle.ntt_forward(0, a)
    for i in (0, 1, 2, 63, 64, 127):
        g = pow(5, i, 2 * N)
        ta = oracle.automorphic_rns(a, g, Q)
        assert np.array_equal(oracle.ntt_forward(0, ta), fa[perms[i]]), i
    assert np.array_equal(perms[0], np.arange(N))


def test_query_encode_is_slot_of_rotation_exponent(oracle):
    enc = oracle.query_encode()
    h = oracle.hexl_ntt_index()
    slot_of = {int(e): s for s, e in enumerate(h)}
    for i in range(N):
        e = pow(5, i, 2 * N) if i < N // 2 else (2 * N - 1) * pow(5, N - 1 - i, 2 * N) % (2 * N)
        assert int(enc[i]) == slot_of[e]


# ----------------------------------------------------------- decompositions

def _signed(v, q):
    v = int(v)
    return v - q if v > q // 2 else v


@pytest.mark.parametrize("ell,base,bg", [(2, 1 << 20, 1 << 18), (4, 0, 1 << 14)])
def test_decompose_reconstructs(oracle, ell, base, bg):
    """utils.cpp:815-867 recontruct: sum_j d_j * PP * Bg^j == x up to < PP/2."""
    Q = oracle.Q
    rng = np.random.default_rng(11)
    x = rng.integers(0, Q, N, dtype=np.uint64)
    bb = base.bit_length() - 1 if base else 0
    gb = bg.bit_length() - 1
    # special values: 0, q-1, the `d > q` boundary x = q - offset, and q itself
    offset = sum(1 << (bb + i * gb) for i in range(ell)) << (gb - 1)
    offset += base // 2
    specials = [0, Q - 1, Q - offset, Q - offset - 1, Q - offset + 1, Q]
    x[:len(specials)] = specials
    d = oracle.decompose(x, ell, base, bg, Q)
    pp = max(base, 1)
    for i in range(N):
        digits = [_signed(d[j, i], Q) for j in range(ell)]
        assert all(-bg // 2 <= v < bg // 2 for v in digits)
        rec = sum(v * pp * bg ** j for j, v in enumerate(digits))
        err = (rec - int(x[i])) % Q
        err = err - Q if err > Q // 2 else err
        assert abs(err) <= max(base // 2, 1) + (1 if i < len(specials) else 0), (i, err)
    # x = q behaves exactly like x = 0 (reference automorphisms emit q for 0)
    assert np.array_equal(d[:, 0], d[:, 5])


def test_decompose_gt_not_ge(oracle):
    """utils.cpp:784 reduces with d > q: at x = q - offset, d == q stays unreduced."""
    Q = oracle.Q
    ell, base, bg = 2, 1 << 20, 1 << 18
    offset = (((1 << 20) + (1 << 38)) << 17) + (1 << 19)
    x = np.zeros(N, dtype=np.uint64)
    x[0] = Q - offset          # d == Q: not reduced
    x[1] = Q - offset + 1      # d == Q + 1: reduced to 1
    d = oracle.decompose(x, ell, base, bg, Q)
    D0 = Q >> 20
    want0 = [((D0 >> (18 * j)) & ((1 << 18) - 1)) - (1 << 17) for j in range(2)]
    assert [_signed(d[j, 0], Q) for j in range(2)] == want0
    assert [_signed(d[j, 1], Q) for j in range(2)] == [-(1 << 17)] * 2


def test_decompose_rlwe_signed(oracle):
    Q = oracle.Q
    rng = np.random.default_rng(5)
    a = rng.integers(0, Q, N, dtype=np.uint64)
    b = rng.integers(0, Q, N, dtype=np.uint64)
    d = oracle.decompose_rlwe(a, b, 2, 1 << 20, 1 << 18, Q)
    for src, base in ((a, 0), (b, 2)):
        for i in range(0, N, 97):
            digits = [_signed(d[base + j, i], Q) for j in range(2)]
            assert all(-(1 << 17) <= v < (1 << 17) for v in digits)
            xs = _signed(src[i], Q) >> 20
            # two signed digits cover the value up to the dropped top carry
            rec = digits[0] + digits[1] * (1 << 18)
            assert (xs - rec) % (1 << 36) == 0


def test_crt_inv_bsgs_and_digits(oracle):
    Q, bs = oracle.Q, oracle.BS
    rng = np.random.default_rng(9)
    xq = rng.integers(0, Q, N, dtype=np.uint64)
    xb = rng.integers(0, bs, N, dtype=np.uint64)
    v = oracle.crt_inv_bsgs(xq, xb)
    for i in range(0, N, 131):
        assert v[i] < Q * bs and v[i] % Q == int(xq[i]) and v[i] % bs == int(xb[i])
    dq, db = oracle.decompose_bsgs(xq, xb)
    for i in range(0, N, 131):
        dig = [_signed(dq[j, i], Q) for j in range(3)]
        assert dig == [_signed(db[j, i], bs) for j in range(3)]
        rec = sum(dv * (1 << 20) * (1 << 20) ** j for j, dv in enumerate(dig))
        err = (rec - v[i]) % (Q * bs)
        err = err - Q * bs if err > Q * bs // 2 else err
        assert abs(err) <= 1 << 19


def test_crt_compose(oracle):
    Q, q1, q2 = oracle.Q, oracle.Q1, oracle.Q2
    rng = np.random.default_rng(1)
    for _ in range(200):
        x, y = (int(v) for v in rng.integers(0, 1 << 63, 2, dtype=np.uint64))
        z = int(oracle.lib().ko_crt_compose(x, y))
        assert z < Q and z % q1 == x % q1 and z % q2 == y % q2
        R = oracle.ref_lib()
        if R is not None:
            assert int(R.ref_crt_compose(x, y)) == z


def test_db_product_matches_reference_kernel(oracle):
    """oracle db_product == the reference's own AVX-512 fastMultiplyQueryByDatabaseDim1InvCRT."""
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (needs /root/reference and an AVX-512 host)")
    n1, n2 = 128, 16
    rng = np.random.default_rng(2)
    q1, q2 = oracle.Q1, oracle.Q2

    def packed(shape):
        lo = rng.integers(0, q1, shape, dtype=np.uint64)
        hi = rng.integers(0, q2, shape, dtype=np.uint64)
        return lo | (hi << np.uint64(32))

    db = packed(N * n2 * n1)
    cbuf = packed(N * n1 * 2)
    got = oracle.db_product(db, cbuf, n1, n2)
    want = oracle.ref_db_product(db, cbuf, n1, n2)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n1,n2", [(128, 16), (256, 8), (4, 512)])
def test_db_product_vector_loop_matches_scalar(oracle, n1, n2):
    """the oracle's AVX-512 database loop (the CPU baseline's) and its
    scalar loop give the same words, including all-(q-1) words where the
    unreduced sums are largest"""
    if not oracle.db_product_vectorised():
        pytest.skip("host without AVX-512: the scalar loop is the only one")
    rng = np.random.default_rng(n1)
    q1, q2 = oracle.Q1, oracle.Q2
    lo = rng.integers(0, q1, (2, N * n2 * n1 + N * n1 * 2), dtype=np.uint64)
    hi = rng.integers(0, q2, (2, N * n2 * n1 + N * n1 * 2), dtype=np.uint64)
    words = lo[0] | (hi[0] << np.uint64(32))
    words[:4096] = np.uint64(q1 - 1) | (np.uint64(q2 - 1) << np.uint64(32))
    db, cbuf = words[:N * n2 * n1].copy(), words[N * n2 * n1:].copy()
    cbuf[:2 * n1] = np.uint64(q1 - 1) | (np.uint64(q2 - 1) << np.uint64(32))
    assert np.array_equal(oracle.db_product(db, cbuf, n1, n2), oracle.db_product_scalar(db, cbuf, n1, n2))


# ------------------------------------------------------------ end to end

def test_readme_target_indices(oracle):
    """README.md:56 prints target_col 177, target_packing 4 for the default seed."""
    # glibc rand() with seed 1: 4096 secret draws, then the two target draws
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    for _ in range(N):
        libc.rand()
    assert (libc.rand() % 2048, libc.rand() % 16) == (177, 4)
    sec = np.zeros(N, dtype=np.uint64)
    q = np.zeros((2, 2, N), dtype=np.uint64)
    sh = oracle.shapes(16, 128)
    arrs = [np.zeros(sh[k], dtype=np.uint64) for k in ("rgsw", "baby", "giant", "packing")]
    rc = oracle.lib().ko_gen_client(16, 128, 1, 177, 4, oracle.p64(sec), oracle.p64(q), *(oracle.p64(a) for a in arrs))
    assert rc == 0
    # the secret is ternary and is the first 4096 draws of the stream
    vals = set(int(v) for v in sec)
    assert vals <= {0, 1, oracle.Q - 1}


@pytest.mark.parametrize("r,n1", [(2, 128), (4, 64), (1, 256)])
def test_kat_small(oracle, r, n1):
    """test-pir flow at reduced r: the planted record decrypts to [1..N]."""
    inst = oracle.gen_instance(r=r, n1=n1, seed=1)
    resp = oracle.answer(inst)
    msg = oracle.decrypt(r, inst.secret, resp)
    assert np.array_equal(msg, np.arange(1, N + 1, dtype=np.uint64))


@pytest.mark.slow
def test_kat_test_pir_default(oracle):
    """test-pir.cpp defaults (r 16, N1 128, seed 1) reproduce README.md:56-65."""
    inst = oracle.gen_instance(r=16, n1=128, seed=1)
    assert inst.target == (177, 4)
    resp = oracle.answer(inst)
    msg = oracle.decrypt(16, inst.secret, resp)
    assert np.array_equal(msg, np.arange(1, N + 1, dtype=np.uint64))


def test_planted_mistake_breaks_kat(oracle):
    """A wrong key (one flipped word in a baby key) must break the KAT, so the
    KAT is sensitive to the key-switch path."""
    inst = oracle.gen_instance(r=2, n1=128, seed=5)
    inst.baby[3, 1, 1, 100] = (int(inst.baby[3, 1, 1, 100]) + 12345) % oracle.Q
    resp = oracle.answer(inst)
    msg = oracle.decrypt(2, inst.secret, resp)
    assert not np.array_equal(msg, np.arange(1, N + 1, dtype=np.uint64))


def test_gen_records_matches_gen_db_and_tocrt(oracle):
    """the records the oracle keeps are the input of its preprocessing: the
    same seed gives the words of gen_db, and database_tobsgsntt of a block
    re-laid as database_tocrt (crt.cpp:859-885) gives that block's words"""
    r, n1 = 2, 256
    n2 = N // 2 // n1
    rec, words = oracle.gen_records(r, n1, seed=3)
    assert np.array_equal(words, oracle.gen_db(r, n1, seed=3))
    assert rec.max() < 65537
    nt = oracle.database_tobsgsntt(n1, rec[1])
    q1, q2 = np.uint64(268369921), np.uint64(249561089)
    packed = (nt % q1) | ((nt % q2) << np.uint64(32))          # [n2*n1][N]
    want = packed.reshape(n2, n1, N).transpose(2, 0, 1).reshape(-1)
    assert np.array_equal(words[1], want)


@pytest.mark.parametrize("n1", [64, 256])
def test_database_tocrt_matches_reference(oracle, n1):
    """the oracle's database words of a block equal the reference's own
    database_tocrt (crt.cpp:859-885, built into oracle/_ref) applied to the
    oracle's database_tobsgsntt of that block's records"""
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rec, words = oracle.gen_records(2, n1, seed=11)
    for b in range(2):
        nt = oracle.database_tobsgsntt(n1, rec[b])
        assert np.array_equal(oracle.ref_database_tocrt(nt, n1), words[b]), b


def test_records_to_words_matches_gen_records(oracle):
    """records_to_words (the oracle's preprocessing of given records) gives
    gen_records' words, and block stride 0 repeats block 0's words"""
    r, n1 = 2, 128
    rec, words = oracle.gen_records(r, n1, seed=5)
    assert np.array_equal(oracle.records_to_words(r, n1, rec), words)
    rep = oracle.records_to_words(r, n1, rec[:1], block_stride=0)
    assert np.array_equal(rep[0], words[0]) and np.array_equal(rep[1], words[0])


@pytest.mark.parametrize("r,n1", [(2, 128), (4, 64)])
def test_planted_block_round_trip(oracle, r, n1):
    """one planted block repeated r times (the shape the full-size GPU
    round trips use): each client recovers m - (m mod r) + 1 + shift; with
    one block and shift 0 at r = 1 this is test-pir's 1..N"""
    from tests.planted import planted_block, planted_message
    cols, packs, shifts = [5, 1000], [r - 1, 0], [0, 9]
    rec = planted_block(oracle, r, cols, packs, shifts)
    words = oracle.records_to_words(r, n1, rec[None], block_stride=0)
    perms = oracle.perm_matrix(n1)
    for c in range(2):
        cl = oracle.gen_client(r, n1, 40 + c, cols[c], packs[c])
        resp = oracle.answer(r=r, n1=n1, db=words, query=cl["query"], rgsw=cl["rgsw"], baby=cl["baby"],
                             giant=cl["giant"], packing=cl["packing"], perms=perms)
        assert np.array_equal(oracle.decrypt(r, cl["secret"], resp), planted_message(r, shifts[c])), c


def test_reference_crtbamod_composition_overflows():
    """why offline key derivation (genAutoKeyFromOffline) is not restated: its
    crtBaMod half composes each coefficient with crt_inv_bsgs_ba
    (utils.cpp:555-571) through compute_mult_qinvqi_bsgs_ba (utils.cpp:434-460),
    whose (temp >> 64) * two128_mod needs up to 142 bits in a 128-bit type.
    The arithmetic restated here with that wrap-around gives a value that is
    not the CRT composition for every sampled input, while the crtMod * auxMod
    half (compute_mult_qinvqi_bsgs_aux, the giant-step keys) stays exact"""
    import random
    q1, q2, bs, aux = 268369921, 249561089, 16760833, 268361729
    Q = q1 * q2
    w = 1 << 128

    def compose(x1, x2, m2, qinv2, qinv1, two128):
        mod = Q * m2
        assert two128 == w % mod

        def term(inp, qinv, qi):
            f = qi * qinv
            a, b = f >> 64, f & ((1 << 64) - 1)
            temp = inp * a % w
            r = ((temp >> 64) * two128 % w) % mod
            r += ((temp << 64) % w) % mod
            r += inp * b % w
            return r % w % mod
        return (term(x1, qinv2, m2) + term(x2, qinv1, Q)) % w % mod

    rng = random.Random(1)
    ba = bs * aux
    ba_two128 = (14764224294473 << 63) + 1483995998376346528         # crt.h:60-61
    aux_two128 = (1379708 << 63) + 2086958912630458341               # crt.h:65-66
    wrong_ba = wrong_aux = 0
    for _ in range(200):
        x1, x2 = rng.randrange(Q), rng.randrange(ba)
        v = compose(x1, x2, ba, 26456327859403970, 2721180491014976, ba_two128)    # crt.h:53-54
        wrong_ba += v % Q != x1 or v % ba != x2
        y2 = rng.randrange(aux)
        u = compose(x1, y2, aux, 48735363086513670, 73083388, aux_two128)          # crt.h:56, 63
        wrong_aux += u % Q != x1 or u % aux != y2
    assert wrong_ba == 200 and wrong_aux == 0


# ---------------------------------------------------------------- offline giant-step key derivation

AUX = 268361729


def _auto_secret(s, g, q):
    out = [0] * N
    for i, v in enumerate(s):
        d = (i * g) % (2 * N)
        if d >= N:
            out[d - N] = (q - int(v)) % q
        else:
            out[d] = int(v)
    return out


def _centered(x, q):
    x %= q
    return x - q if x > q // 2 else x


def test_aux_constants_and_root(oracle):
    """crt.h:37,56,63 constants and the auxMod NTT root (lwe.cpp:1016-1019,
    HEXL's minimal primitive 2N-th root)"""
    Q = oracle.Q
    assert oracle.const(10) == AUX
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):   # deterministic Miller-Rabin below 3.3e24
        d, s = AUX - 1, 0
        while d % 2 == 0:
            d, s = d // 2, s + 1
        x = pow(a, d, AUX)
        assert x in (1, AUX - 1) or any(pow(x, 2 ** t, AUX) == AUX - 1 for t in range(1, s))
    assert oracle.const(11) * AUX % Q == 1
    assert oracle.const(12) * Q % AUX == 1
    root = oracle.const(13)
    assert pow(root, N, AUX) == AUX - 1
    assert all(pow(x, N, AUX) != AUX - 1 for x in range(2, root))


def test_crt_inv_bsgs_aux_is_the_crt_composition(oracle):
    """utils.cpp:535-553 with utils.cpp:462-488: exact CRT into [0, crtMod*auxMod),
    also for the non-canonical zeros q - 0 an automorphism writes"""
    Q = oracle.Q
    rng = np.random.default_rng(3)
    xq = rng.integers(0, Q, N, dtype=np.uint64)
    xa = rng.integers(0, AUX, N, dtype=np.uint64)
    xq[:4] = [0, Q, Q - 1, 0]
    xa[:4] = [0, AUX, 0, AUX - 1]
    got = oracle.crt_inv_bsgs_aux(xq, xa)
    M = Q * AUX
    for i in range(N):
        v = got[i]
        assert 0 <= v < M and v % Q == int(xq[i]) % Q and v % AUX == int(xa[i]) % AUX, i


def test_decompose_bsgs_aux_recombines(oracle):
    """utils.cpp:1191-1256 (ell 5, PP = Bg = 2^14): signed digits in
    [-2^13, 2^13) whose gadget sum is the composed value to within PP/2 (mod M)"""
    Q = oracle.Q
    M = Q * AUX
    rng = np.random.default_rng(4)
    xq = rng.integers(0, Q, N, dtype=np.uint64)
    xa = rng.integers(0, AUX, N, dtype=np.uint64)
    xq[:3] = [0, Q - 1, 0]
    xa[:3] = [0, AUX - 1, 0]
    dq, da = oracle.decompose_bsgs_aux(xq, xa)
    comp = oracle.crt_inv_bsgs_aux(xq, xa)
    for i in range(N):
        s = 0
        for j in range(5):
            r = _centered(int(dq[j, i]), Q)
            assert -(1 << 13) <= r < (1 << 13)
            assert int(da[j, i]) == r % AUX
            s += r << (14 + 14 * j)
        assert abs(_centered(s - comp[i], M)) <= 1 << 13, i


@pytest.mark.parametrize("n1", [128, 1024])
def test_offline_giant_keys_decrypt(oracle, n1):
    """twosteps.cpp:1814-1863: every key of the derivation chain (5^(k N1),
    k >= 2) decrypts under the secret to tau(s) * 2^(11+15j) * auxMod plus a
    noise that is the same integer in its crtMod and auxMod parts; every
    switched key for k >= 2 decrypts mod crtMod to tau(s) * 2^(11+15j) plus a
    small noise.  The key at k = 1 is switched from the stored NTT-form words,
    as the reference does, and does not decrypt."""
    Q = oracle.Q
    s, off, auto = oracle.gen_offline_giant(n1, 5)
    chain, giant = oracle.giant_from_offline(n1, off, auto)
    n2 = N // 2 // n1
    assert chain.shape[0] == n2 - 2 and giant.shape[0] == n2 - 1
    for k in range(2, n2):
        ts = _auto_secret(s, pow(5, k * n1, 2 * N), Q)
        for j in range(3):
            m = (1 << (11 + 15 * j)) * AUX
            e1 = oracle.rlwe_phase(0, s, chain[k - 2, j, 0], chain[k - 2, j, 1], False)
            e2 = oracle.rlwe_phase(5, s, chain[k - 2, j + 3, 0], chain[k - 2, j + 3, 1], False)
            eg = oracle.rlwe_phase(0, s, giant[k - 1, j, 0], giant[k - 1, j, 1], True)
            for i in range(0, N, 5):
                e = _centered(int(e1[i]) - ts[i] * (m % Q), Q)
                assert abs(e) < 1 << 32 and e % AUX == int(e2[i]), (k, j, i)
                assert abs(_centered(int(eg[i]) - ts[i] * (1 << (11 + 15 * j)), Q)) < 1 << 16, (k, j, i)
    ts = _auto_secret(s, pow(5, n1, 2 * N), Q)
    e0 = oracle.rlwe_phase(0, s, giant[0, 0, 0], giant[0, 0, 1], True)
    assert max(abs(_centered(int(e0[i]) - ts[i] * (1 << 11), Q)) for i in range(N)) > 1 << 40
