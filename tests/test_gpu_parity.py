"""GPU parity: the HIP path (through the C ABI) against the CPU oracle, word
for word, on the same seeded inputs.  Exact integer arithmetic: tolerance 0."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096
_cache = {}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import kspir_amd
    c = kspir_amd.Context(0)
    yield c
    c.close()


def instance(oracle, r, n1, seed):
    key = (r, n1, seed)
    if key not in _cache:
        inst = oracle.gen_instance(r=r, n1=n1, seed=seed)
        resp, st = oracle.answer(inst, want_stages=True)
        _cache.clear()
        _cache[key] = (inst, resp, st)
    return _cache[key]


def upload(ctx, inst):
    db = ctx.database(inst.db, inst.r, inst.n1)
    keys = ctx.keys(inst.r, inst.n1, inst.baby, inst.giant, inst.packing, inst.rgsw, inst.perms)
    return db, keys


@pytest.mark.parametrize("prime,modsel,q", [(0, 3, 268369921), (1, 4, 249561089), (2, 1, 16760833),
                                             (4, 5, 268361729)])
def test_ntt_matches_oracle(ctx, oracle, prime, modsel, q):
    say(f"[gpu] ntt prime={prime}")
    rng = np.random.default_rng(prime)
    polys = rng.integers(0, q, (4, N), dtype=np.uint32)
    polys[0, :] = 0
    polys[1, :] = q - 1
    fwd = ctx.test_ntt(polys, prime, inverse=False)
    for i in range(4):
        assert np.array_equal(fwd[i].astype(np.uint64), oracle.ntt_forward(modsel, polys[i].astype(np.uint64))), i
    back = ctx.test_ntt(fwd, prime, inverse=True)
    assert np.array_equal(back, polys)


def test_stages_r2(ctx, oracle):
    """every stage of one answer at r 2, N1 128 against the oracle's"""
    say("[gpu] stages r=2 n1=128 seed=3")
    inst, resp, st = instance(oracle, 2, 128, 3)
    db, keys = upload(ctx, inst)
    got, gst = ctx.answer_inspect(db, keys, inst.query)
    assert np.array_equal(gst["cipherbuf"], st["cipherbuf"]), "rotated query words (reorientCipher)"
    assert np.array_equal(gst["firstdim"], st["firstdim"]), "first-dimension result"
    assert np.array_equal(gst["ext"], st["ext"]), "external product"
    assert np.array_equal(got, resp), "response words"
    msg = oracle.decrypt(2, inst.secret, got)
    assert np.array_equal(msg, np.arange(1, N + 1, dtype=np.uint64))
    db.close()
    keys.close()


@pytest.mark.parametrize("r,n1,seed", [(1, 256, 2), (4, 64, 4), (8, 512, 6), (16, 128, 1)])
def test_answer_bit_exact(ctx, oracle, r, n1, seed):
    """response words identical to the oracle's; planted record decrypts to [1..N]"""
    say(f"[gpu] answer r={r} n1={n1} seed={seed}")
    inst, resp, _ = instance(oracle, r, n1, seed)
    db, keys = upload(ctx, inst)
    got = ctx.answer(db, keys, inst.query)
    assert np.array_equal(got, resp)
    msg = oracle.decrypt(r, inst.secret, got)
    assert np.array_equal(msg, np.arange(1, N + 1, dtype=np.uint64))
    # repeated calls on the resident handles give the same words
    assert np.array_equal(ctx.answer(db, keys, inst.query), resp)
    db.close()
    keys.close()


def test_stage_entry_points(ctx, oracle):
    """matrix_vector_mul_bsgs_rns_crt_large / externalProduct / packingRLWEs
    called one at a time, as test-pir.cpp:111-115 does"""
    say("[gpu] stage entry points r=4 n1=128 seed=8")
    inst, resp, st = instance(oracle, 4, 128, 8)
    db, keys = upload(ctx, inst)
    res, q_after = ctx.matrix_vector_mul(db, keys, inst.query)
    assert np.array_equal(res, st["firstdim"])
    # input[0] becomes the modulus-switched query in NTT form (twosteps.cpp:1074-1086)
    cb = st["cipherbuf"][:, 0, :]
    lo, hi = (cb & np.uint64(0xffffffff)), (cb >> np.uint64(32))
    want = np.array([[int(oracle.lib().ko_crt_compose(int(lo[z, ab]), int(hi[z, ab]))) for z in range(N)]
                     for ab in range(2)], dtype=np.uint64)
    assert np.array_equal(q_after[0], want)
    assert np.array_equal(q_after[1], inst.query[1])
    ext = []
    for j in range(inst.r):
        e, fd_coef = ctx.external_product(keys, np.zeros((2, N), dtype=np.uint64), res[j], True)
        assert np.array_equal(e, st["ext"][j]), j
        # like the reference, the first-dimension input comes back in coefficient form
        assert np.array_equal(fd_coef[0], oracle.ntt_inverse(0, res[j][0]))
        assert np.array_equal(fd_coef[1], oracle.ntt_inverse(0, res[j][1]))
        ext.append(e)
    out = ctx.packing_rlwes(keys, np.stack(ext))
    assert np.array_equal(out, resp)
    db.close()
    keys.close()


@pytest.mark.parametrize("waves", [2, 8, 16])
def test_db_pass_workgroup_shapes(ctx, oracle, waves):
    """every workgroup shape of the single-query database pass gives the
    oracle's response (r 4, N1 64: N2 32 rows over 2..16 waves)"""
    say(f"[gpu] db_waves={waves} r=4 n1=64 seed=4")
    inst, resp, _ = instance(oracle, 4, 64, 4)
    db, keys = upload(ctx, inst)
    ctx.set("db_waves", waves)
    try:
        assert np.array_equal(ctx.answer(db, keys, inst.query), resp)
    finally:
        ctx.set("db_waves", 4)
    db.close()
    keys.close()


def test_keys_and_db_shape_checks(ctx, oracle):
    """mismatched handles and bad shapes are rejected with an error, not run"""
    import kspir_amd
    say("[gpu] shape checks")
    inst, _, _ = instance(oracle, 4, 128, 8)
    with pytest.raises(kspir_amd.KspirError):
        ctx.database(inst.db, 3, 128)
    with pytest.raises(kspir_amd.KspirError):
        ctx.database(inst.db, 4, 96)
    bad = inst.perms.copy()
    bad[1, 5] = N
    with pytest.raises(kspir_amd.KspirError):
        ctx.keys(4, 128, inst.baby, inst.giant, inst.packing, inst.rgsw, bad)
    dup = inst.perms.copy()
    dup[3, 7] = dup[3, 8]      # in range, but row 3 no longer a permutation
    with pytest.raises(kspir_amd.KspirError):
        ctx.keys(4, 128, inst.baby, inst.giant, inst.packing, inst.rgsw, dup)
    db = ctx.database(inst.db[:2].copy(), 2, 128)
    keys = ctx.keys(4, 128, inst.baby, inst.giant, inst.packing, inst.rgsw, inst.perms)
    with pytest.raises(kspir_amd.KspirError):
        ctx.answer(db, keys, inst.query)
    db.close()
    keys.close()


def test_batch_matches_single_and_oracle(ctx, oracle):
    """a batch of queries from different clients (own secrets and keys, seeds
    11..16, random targets) gives, for every query, the words the query
    gets alone, and those equal the oracle's"""
    say("[gpu] batch r=4 n1=128 nq=6")
    r, n1 = 4, 128
    db_words = oracle.gen_db(r, n1, seed=99)
    perms = oracle.perm_matrix(n1)
    rng = np.random.default_rng(0)
    clients = []
    for c in range(6):
        tcol, tpack = int(rng.integers(0, N // 2)), int(rng.integers(0, r))
        clients.append(oracle.gen_client(r, n1, 11 + c, tcol, tpack))
    db = ctx.database(db_words, r, n1)
    keys = [ctx.keys(r, n1, cl["baby"], cl["giant"], cl["packing"], cl["rgsw"], perms) for cl in clients]
    queries = np.stack([cl["query"] for cl in clients])
    got = ctx.answer_batch(db, keys, queries)
    for c, cl in enumerate(clients):
        alone = ctx.answer(db, keys[c], cl["query"])
        assert np.array_equal(got[c], alone), c
        want = oracle.answer(r=r, n1=n1, db=db_words, query=cl["query"], rgsw=cl["rgsw"], baby=cl["baby"],
                             giant=cl["giant"], packing=cl["packing"], perms=perms)
        assert np.array_equal(got[c], want), c
    # the same client twice in one batch, and a batch of one
    got2 = ctx.answer_batch(db, [keys[1], keys[1]], np.stack([queries[1], queries[1]]))
    assert np.array_equal(got2[0], got[1]) and np.array_equal(got2[1], got[1])
    assert np.array_equal(ctx.answer_batch(db, [keys[3]], queries[3:4])[0], got[3])
    for k in keys:
        k.close()
    db.close()


@pytest.mark.parametrize("n1,chunk", [(64, 3), (256, 1 << 30), (128, 2)])
def test_batch_mfma_shapes_and_chunks(ctx, oracle, n1, chunk):
    """the int8-MFMA database pass at N1 64/128/256, with the batch split into
    chunks of `chunk` queries, equals the single-query VALU pass"""
    say(f"[gpu] batch mfma n1={n1} chunk={chunk}")
    r = 4
    db_words = oracle.gen_db(r, n1, seed=7)
    perms = oracle.perm_matrix(n1)
    clients = [oracle.gen_client(r, n1, 31 + c, (37 * c) % (N // 2), c % r) for c in range(5)]
    db = ctx.database(db_words, r, n1)
    keys = [ctx.keys(r, n1, cl["baby"], cl["giant"], cl["packing"], cl["rgsw"], perms) for cl in clients]
    queries = np.stack([cl["query"] for cl in clients])
    ctx.set("mfma_max_chunk", chunk)
    try:
        got = ctx.answer_batch(db, keys, queries)
    finally:
        ctx.set("mfma_max_chunk", 1 << 30)
    for c, cl in enumerate(clients):
        assert np.array_equal(got[c], ctx.answer(db, keys[c], cl["query"])), c
    want0 = oracle.answer(r=r, n1=n1, db=db_words, query=clients[0]["query"], rgsw=clients[0]["rgsw"],
                          baby=clients[0]["baby"], giant=clients[0]["giant"], packing=clients[0]["packing"],
                          perms=perms)
    assert np.array_equal(got[0], want0)
    for k in keys:
        k.close()
    db.close()


@pytest.mark.parametrize("r,n1,nq", [(4, 128, 5), (2, 64, 3), (4, 256, 4), (1, 128, 2)])
def test_batch_mfma_direct_matches_tile_major(ctx, oracle, r, n1, nq):
    """the MFMA database pass that writes its products straight into the
    first-dimension layout (option mfma_direct; two row tiles per wave where
    r*N2 allows, one otherwise) gives the words of the tile-major pass and
    the oracle's"""
    say(f"[gpu] mfma direct r={r} n1={n1} nq={nq}")
    db_words = oracle.gen_db(r, n1, seed=13)
    perms = oracle.perm_matrix(n1)
    clients = [oracle.gen_client(r, n1, 131 + c, (41 * c + 9) % (N // 2), c % r) for c in range(nq)]
    db = ctx.database(db_words, r, n1)
    keys = [ctx.keys(r, n1, cl["baby"], cl["giant"], cl["packing"], cl["rgsw"], perms) for cl in clients]
    queries = np.stack([cl["query"] for cl in clients])
    prev = ctx.get("mfma_direct")
    try:
        ctx.set("mfma_direct", 0)
        tile = ctx.answer_batch(db, keys, queries)
        ctx.set("mfma_direct", 1)
        direct = ctx.answer_batch(db, keys, queries)
    finally:
        ctx.set("mfma_direct", prev)
    assert np.array_equal(direct, tile)
    want0 = oracle.answer(r=r, n1=n1, db=db_words, query=clients[0]["query"], rgsw=clients[0]["rgsw"],
                          baby=clients[0]["baby"], giant=clients[0]["giant"], packing=clients[0]["packing"],
                          perms=perms)
    assert np.array_equal(direct[0], want0)
    for k in keys:
        k.close()
    db.close()


@pytest.mark.parametrize("r,n1,nq", [(4, 128, 2), (4, 64, 5), (2, 512, 3), (8, 256, 1)])
def test_fused_giant_steps(ctx, oracle, r, n1, nq):
    """the fused per-(query, block) giant-step kernel (forced on with
    fused_min_batch = 1) gives the words of the per-row kernels for every
    query, and query 0's equal the oracle's; r*nq = 8 takes the XCD-dealt
    workgroup order, the other shapes the plain one"""
    say(f"[gpu] fused giant steps r={r} n1={n1} nq={nq}")
    db_words = oracle.gen_db(r, n1, seed=5)
    perms = oracle.perm_matrix(n1)
    clients = [oracle.gen_client(r, n1, 51 + c, (101 * c + 3) % (N // 2), c % r) for c in range(nq)]
    db = ctx.database(db_words, r, n1)
    keys = [ctx.keys(r, n1, cl["baby"], cl["giant"], cl["packing"], cl["rgsw"], perms) for cl in clients]
    queries = np.stack([cl["query"] for cl in clients])
    plain = ctx.answer_batch(db, keys, queries)
    ctx.set("fused_min_batch", 1)
    try:
        fused = ctx.answer_batch(db, keys, queries)
        ctx.set("giant_dma", 0)            # rows loaded through registers
        fused_regs = ctx.answer_batch(db, keys, queries)
        ctx.set("giant_wgs_per_cu", 2)     # the four-plane, two-per-CU variant
        fused2 = ctx.answer_batch(db, keys, queries)
    finally:
        ctx.set("fused_min_batch", 16)
        ctx.set("giant_wgs_per_cu", 1)
        ctx.set("giant_dma", 1)
    assert np.array_equal(fused, plain)
    assert np.array_equal(fused_regs, plain)
    assert np.array_equal(fused2, plain)
    want0 = oracle.answer(r=r, n1=n1, db=db_words, query=clients[0]["query"], rgsw=clients[0]["rgsw"],
                          baby=clients[0]["baby"], giant=clients[0]["giant"], packing=clients[0]["packing"],
                          perms=perms)
    assert np.array_equal(fused[0], want0)
    for k in keys:
        k.close()
    db.close()


@pytest.mark.parametrize("r,n1", [(2, 128), (1, 64), (2, 2048)])
def test_database_preprocessing_on_gpu(ctx, oracle, r, n1):
    """database_tobsgsntt + database_tocrt on the GPU (records -> resident
    words) equal the oracle's words, export returns them in the reference
    layout, and an export loads back to the same words; block stride 0
    repeats block 0"""
    say(f"[gpu] database preprocessing r={r} n1={n1}")
    rec, want = oracle.gen_records(r, n1, seed=17, plant=True, tcol=5 % (N // 2), tpack=r - 1)
    db = ctx.database_from_records(rec, r, n1)
    got = db.export()
    assert np.array_equal(got, want)
    db2 = ctx.database(got, r, n1)
    assert np.array_equal(db2.export(), want)
    db3 = ctx.database_from_records(rec, r, n1, block_stride=0)
    rep = db3.export()
    assert all(np.array_equal(rep[b], want[0]) for b in range(r))
    one = ctx.database_tobsgsntt(rec[0], n1)
    assert np.array_equal(one, oracle.database_tobsgsntt(n1, rec[0]))
    for d in (db, db2, db3):
        d.close()


@pytest.mark.parametrize("lanes", [1, 3])
def test_serving_loop_mixed_databases(ctx, oracle, lanes):
    """queries for two databases submitted interleaved to the serving loop
    (batch cap 3, a budget that admits two queries of the larger shape per
    lane; 1 or 3 lanes answering batches at once, the extra lanes on contexts
    of their own) come back equal to the same queries answered alone, and
    equal to the oracle"""
    import kspir_amd
    say(f"[gpu] serving loop, two databases, lanes={lanes}")
    shapes = [(2, 128, 61), (4, 64, 62)]
    dbs, clients, want = [], [], []
    for r, n1, seed in shapes:
        dbw = oracle.gen_db(r, n1, seed=seed)
        perms = oracle.perm_matrix(n1)
        db = ctx.database(dbw, r, n1)
        cl = [oracle.gen_client(r, n1, seed * 10 + c, (13 * c) % (N // 2), c % r) for c in range(4)]
        ks = [ctx.keys(r, n1, x["baby"], x["giant"], x["packing"], x["rgsw"], perms) for x in cl]
        dbs.append((db, dbw, r, n1, perms))
        clients.append((cl, ks))
        want.append([ctx.answer(db, k, x["query"]) for x, k in zip(cl, ks)])
    w0 = oracle.answer(r=shapes[0][0], n1=shapes[0][1], db=dbs[0][1], query=clients[0][0][0]["query"],
                       rgsw=clients[0][0][0]["rgsw"], baby=clients[0][0][0]["baby"], giant=clients[0][0][0]["giant"],
                       packing=clients[0][0][0]["packing"], perms=dbs[0][4])
    assert np.array_equal(want[0][0], w0)
    foot = max(kspir_amd.query_footprint(r, n1) for r, n1, _ in shapes)
    srv = kspir_amd.Server(ctx, max_batch=3, budget_bytes=2 * foot * lanes, lanes=lanes)
    try:
        tickets = []
        for c in range(4):
            for g in range(2):
                cl, ks = clients[g]
                tickets.append((g, c, srv.submit(dbs[g][0], ks[c], cl[c]["query"])))
        for g, c, t in tickets:
            resp, ms = srv.wait(t)
            assert ms > 0
            assert np.array_equal(resp, want[g][c]), (g, c)
        st = srv.stats()
        assert st["answered"] == 8 and st["pending"] == 0 and st["batches"] >= 4
    finally:
        srv.close()
    for db, *_ in dbs:
        db.close()
    for cl, ks in clients:
        for k in ks:
            k.close()


@pytest.mark.parametrize("lanes", [1, 2])
def test_serving_loop_batches_span_databases(ctx, oracle, lanes):
    """queries for three databases of one N1 (r 2, 8, 4) submitted interleaved
    to the serving loop come back equal to the same queries answered alone
    (and to the oracle for one of them), with at least one batch spanning
    databases (one launch per stage: once the oldest waiting query is one
    of the r=8 database's, the smaller databases' queries join its batch);
    the same submissions with the option serve_multi off never span"""
    import kspir_amd
    say(f"[gpu] serving loop, batches spanning three databases, lanes={lanes}")
    n1 = 128
    perms = oracle.perm_matrix(n1)
    dbs, clients, want = [], [], []
    for g, (r, seed) in enumerate(((2, 71), (8, 72), (4, 73))):
        dbw = oracle.gen_db(r, n1, seed=seed)
        db = ctx.database(dbw, r, n1)
        cl = [oracle.gen_client(r, n1, seed * 10 + c, (29 * c + 5) % (N // 2), c % r) for c in range(4)]
        ks = [ctx.keys(r, n1, x["baby"], x["giant"], x["packing"], x["rgsw"], perms) for x in cl]
        dbs.append((db, dbw, r))
        clients.append((cl, ks))
        want.append([ctx.answer(db, k, x["query"]) for x, k in zip(cl, ks)])
    db1, dbw1, r1 = dbs[1]
    x = clients[1][0][3]
    w = oracle.answer(r=r1, n1=n1, db=dbw1, query=x["query"], rgsw=x["rgsw"], baby=x["baby"], giant=x["giant"],
                      packing=x["packing"], perms=perms)
    assert np.array_equal(want[1][3], w)
    try:
        for multi in (1, 0):
            ctx.set("serve_multi", multi)
            # a cap above one database's 16 queries: a batch that starts
            # with the r=8 database's queries has room for the smaller ones
            srv = kspir_amd.Server(ctx, max_batch=64, lanes=lanes)
            try:
                # four rounds of the twelve queries, so that batches form
                # while earlier ones run whatever the relative speeds
                tickets = []
                for rnd in range(4):
                    for c in range(4):
                        for g in range(3):
                            cl, ks = clients[g]
                            tickets.append((g, c, srv.submit(dbs[g][0], ks[c], cl[c]["query"])))
                for g, c, t in tickets:
                    resp, ms = srv.wait(t)
                    assert ms > 0
                    assert np.array_equal(resp, want[g][c]), (multi, g, c)
                st = srv.stats()
                say(f"[gpu] serve_multi={multi}: {st}")
                assert st["answered"] == 48 and st["pending"] == 0
                assert (st["spanning"] >= 1) if multi else (st["spanning"] == 0)
            finally:
                srv.close()
    finally:
        ctx.set("serve_multi", 0)
        for db, *_ in dbs:
            db.close()
        for cl, ks in clients:
            for k in ks:
                k.close()


def test_concurrent_contexts_uneven_databases(ctx, oracle):
    """batches against two databases of uneven size (r 2 and 8), each
    enqueued on a context of its own before either is waited for, so they
    overlap on the device, come back equal to the same batches answered one
    after the other on one context, and equal to the oracle"""
    import kspir_amd
    say("[gpu] concurrent contexts, databases r=2 and r=8")
    shapes = [(8, 128, 81, 17), (2, 64, 82, 3)]   # r, n1, seed, queries (17: the fused giant path)
    groups = []
    for r, n1, seed, nq in shapes:
        dbw = oracle.gen_db(r, n1, seed=seed)
        perms = oracle.perm_matrix(n1)
        db = ctx.database(dbw, r, n1)
        cl = [oracle.gen_client(r, n1, seed * 10 + c, (29 * c) % (N // 2), c % r) for c in range(min(nq, 3))]
        cl = [cl[c % len(cl)] for c in range(nq)]
        ks = [ctx.keys(r, n1, x["baby"], x["giant"], x["packing"], x["rgsw"], perms) for x in cl[:3]]
        ks = [ks[c % len(ks)] for c in range(nq)]
        q = np.stack([x["query"] for x in cl])
        qb, rb = ctx.alloc(nq * 4 * N * 8), ctx.alloc(nq * 2 * N * 8)
        qb.upload(q)
        groups.append((r, n1, nq, dbw, perms, db, cl, ks, qb, rb))
    want = []
    for r, n1, nq, dbw, perms, db, cl, ks, qb, rb in groups:
        want.append(ctx.answer_batch(db, ks, np.stack([x["query"] for x in cl])))
    other = kspir_amd.Context(0)
    try:
        for _ in range(2):
            for c, (r, n1, nq, dbw, perms, db, cl, ks, qb, rb) in zip((ctx, other), groups):
                rb.upload(np.zeros((nq, 2, N), dtype=np.uint64))
            for c, (r, n1, nq, dbw, perms, db, cl, ks, qb, rb) in zip((ctx, other), groups):
                c.answer_batch_device(db, ks, qb, rb)
            ctx.synchronize()
            other.synchronize()
            for (r, n1, nq, dbw, perms, db, cl, ks, qb, rb), w in zip(groups, want):
                assert np.array_equal(rb.download(np.zeros((nq, 2, N), dtype=np.uint64)), w), r
    finally:
        other.close()
    r, n1, nq, dbw, perms, db, cl, ks, qb, rb = groups[1]
    w0 = oracle.answer(r=r, n1=n1, db=dbw, query=cl[0]["query"], rgsw=cl[0]["rgsw"], baby=cl[0]["baby"],
                       giant=cl[0]["giant"], packing=cl[0]["packing"], perms=perms)
    assert np.array_equal(want[1][0], w0)
    for r, n1, nq, dbw, perms, db, cl, ks, qb, rb in groups:
        for k in ks[:3]:
            k.close()
        db.close()
        qb.free()
        rb.free()


def test_packing_paths_agree(ctx, oracle):
    """the fused packing level (default) and the three-kernel level give the
    same response words, single and batched, and equal the oracle"""
    say("[gpu] packing paths r=8 n1=128")
    r, n1 = 8, 128
    db_words = oracle.gen_db(r, n1, seed=71)
    perms = oracle.perm_matrix(n1)
    cl = [oracle.gen_client(r, n1, 700 + c, (29 * c + 1) % (N // 2), (3 * c) % r) for c in range(3)]
    db = ctx.database(db_words, r, n1)
    keys = [ctx.keys(r, n1, x["baby"], x["giant"], x["packing"], x["rgsw"], perms) for x in cl]
    queries = np.stack([x["query"] for x in cl])
    ctx.set("fused_min_batch", 1)      # the fused levels run from one query up
    try:
        fused_b = ctx.answer_batch(db, keys, queries)
        fused_1 = ctx.answer(db, keys[0], cl[0]["query"])
        # levels i = 4, 2, 1 carry 12, 6, 3 (pair, query) workgroups: the first
        # two fused, the last as three kernels
        ctx.set("pack_fused_min_wgs", 5)
        mixed_b = ctx.answer_batch(db, keys, queries)
        ctx.set("pack_fused_min_wgs", 0)
        ctx.set("fused_pack", 0)
        plain_b = ctx.answer_batch(db, keys, queries)
        plain_1 = ctx.answer(db, keys[0], cl[0]["query"])
    finally:
        ctx.set("fused_pack", 1)
        ctx.set("pack_fused_min_wgs", 0)
        ctx.set("fused_min_batch", 16)
    assert np.array_equal(fused_b, plain_b)
    assert np.array_equal(mixed_b, plain_b)
    assert np.array_equal(fused_1, plain_1) and np.array_equal(fused_1, fused_b[0])
    want = oracle.answer(r=r, n1=n1, db=db_words, query=cl[0]["query"], rgsw=cl[0]["rgsw"], baby=cl[0]["baby"],
                         giant=cl[0]["giant"], packing=cl[0]["packing"], perms=perms)
    assert np.array_equal(fused_1, want)
    for k in keys:
        k.close()
    db.close()


def test_recover_on_gpu(ctx, oracle):
    """decrypt_bsgs_total on the GPU equals the oracle's messages word for
    word, for the test-pir default and for a batch of clients at r 4"""
    say("[gpu] recover")
    inst, resp, _ = instance(oracle, 16, 128, 1)
    got = ctx.decrypt_bsgs_total(16, inst.secret[None], resp[None])[0]
    assert np.array_equal(got, oracle.decrypt(16, inst.secret, resp))
    assert np.array_equal(got, np.arange(1, N + 1, dtype=np.uint64))
    r, n1 = 4, 64
    db_words = oracle.gen_db(r, n1, seed=81)
    perms = oracle.perm_matrix(n1)
    cl = [oracle.gen_client(r, n1, 810 + c, (11 * c) % (N // 2), c % r) for c in range(3)]
    resps = np.stack([oracle.answer(r=r, n1=n1, db=db_words, query=x["query"], rgsw=x["rgsw"], baby=x["baby"],
                                    giant=x["giant"], packing=x["packing"], perms=perms) for x in cl])
    secrets = np.stack([x["secret"] for x in cl])
    got = ctx.decrypt_bsgs_total(r, secrets, resps)
    for c in range(3):
        assert np.array_equal(got[c], oracle.decrypt(r, secrets[c], resps[c])), c


@pytest.mark.parametrize("n1,seeds", [(64, (31, 32)), (512, (33, 34, 35)), (1024, (36,))])
def test_offline_giant_keys_match_oracle(ctx, oracle, n1, seeds):
    """giant-step half of genAutoKeyFromOffline on the GPU, for several
    clients in one call: every chain key (coefficient form, both RNS parts) and
    every switched giant key (NTT form mod crtMod) equals the oracle's word for
    word; the clients' results do not depend on being batched"""
    say(f"[gpu] offline giant keys n1={n1} clients={len(seeds)}")
    gens = [oracle.gen_offline_giant(n1, s) for s in seeds]
    off = np.stack([g[1] for g in gens])
    auto = np.stack([g[2] for g in gens])
    chain, giant = ctx.gen_giant_keys_from_offline(n1, off, auto)
    for c in range(len(seeds)):
        wc, wg = oracle.giant_from_offline(n1, off[c], auto[c])
        assert np.array_equal(giant[c], wg), c
        if wc is not None and wc.size:
            assert np.array_equal(chain[c], wc), c
    _, g1 = ctx.gen_giant_keys_from_offline(n1, off[-1:], auto[-1:], want_chain=False)
    assert np.array_equal(g1[0], giant[-1])
    import kspir_amd
    with pytest.raises(kspir_amd.KspirError, match="N1"):
        ctx.gen_giant_keys_from_offline(2048, off[:1], auto[:1])   # N2 = 1: no giant steps


def test_graph_replays_match_eager(ctx, oracle):
    """the captured pipeline, replayed four times on the same device buffers
    with new queries and key sets each time and eager answers (the
    inspection path, never captured) in between, gives every replay the words
    of the kernel-by-kernel launch; single-query replays likewise"""
    say("[gpu] graph replays r=4 n1=128 nq=3")
    r, n1 = 4, 128
    db_words = oracle.gen_db(r, n1, seed=91)
    perms = oracle.perm_matrix(n1)
    cl = [oracle.gen_client(r, n1, 900 + c, (53 * c + 7) % (N // 2), c % r) for c in range(5)]
    db = ctx.database(db_words, r, n1)
    keys = [ctx.keys(r, n1, x["baby"], x["giant"], x["packing"], x["rgsw"], perms) for x in cl]
    queries = np.stack([x["query"] for x in cl])
    ctx.set("graphs", 0)
    try:
        eager = np.stack([ctx.answer(db, keys[c], queries[c]) for c in range(5)])
    finally:
        ctx.set("graphs", 1)
    want0 = oracle.answer(r=r, n1=n1, db=db_words, query=cl[0]["query"], rgsw=cl[0]["rgsw"], baby=cl[0]["baby"],
                          giant=cl[0]["giant"], packing=cl[0]["packing"], perms=perms)
    assert np.array_equal(eager[0], want0)
    nq = 3
    qbuf, rbuf = ctx.alloc(nq * 4 * N * 8), ctx.alloc(nq * 2 * N * 8)
    q1buf, r1buf = ctx.alloc(4 * N * 8), ctx.alloc(2 * N * 8)
    for rep in range(4):
        pick = [(rep + j) % 5 for j in range(nq)]
        qbuf.upload(queries[pick])
        ctx.answer_batch_device(db, [keys[c] for c in pick], qbuf, rbuf)
        ctx.synchronize()
        got = rbuf.download(np.zeros((nq, 2, N), dtype=np.uint64))
        for j, c in enumerate(pick):
            assert np.array_equal(got[j], eager[c]), (rep, j, c)
        one = (2 * rep + 1) % 5
        q1buf.upload(queries[one])
        ctx.answer_device(db, keys[one], q1buf, r1buf)
        ctx.synchronize()
        assert np.array_equal(r1buf.download(np.zeros((2, N), dtype=np.uint64)), eager[one]), (rep, one)
        # eager work on the same context between replays
        other = (rep + 3) % 5
        resp, _ = ctx.answer_inspect(db, keys[other], queries[other])
        assert np.array_equal(resp, eager[other]), (rep, other)
    for b in (qbuf, rbuf, q1buf, r1buf):
        b.free()
    for k in keys:
        k.close()
    db.close()


def test_autotune_sets_options_and_answers_stay_exact(ctx, oracle):
    """the autotuner times every path for one database shape, leaves valid
    options behind and a report of its timings; answers under the tuned
    options still equal the oracle's and the single-query answers"""
    say("[gpu] autotune r=4 n1=128 max_batch=8")
    r, n1 = 4, 128
    db_words = oracle.gen_db(r, n1, seed=61)
    perms = oracle.perm_matrix(n1)
    cl = [oracle.gen_client(r, n1, 610 + c, (17 * c + 3) % (N // 2), c % r) for c in range(3)]
    db = ctx.database(db_words, r, n1)
    keys = [ctx.keys(r, n1, x["baby"], x["giant"], x["packing"], x["rgsw"], perms) for x in cl]
    try:
        sb, rep = ctx.autotune(db, keys, max_batch=8)
        assert 1 <= sb <= 8 and rep["serve_batch"] == sb
        assert set(rep["db_waves_ms"]) == {"2", "4", "8", "16"} and ctx.get("db_waves") == rep["db_waves"]
        assert ctx.get("mfma_min_batch") == rep["mfma_min_batch"]
        assert ctx.get("fused_min_batch") == rep["fused_min_batch"]
        assert all(v > 0 for v in rep["step_ms"].values())
        pick = [0, 1, 2, 0, 1]
        got = ctx.answer_batch(db, [keys[c] for c in pick], np.stack([cl[c]["query"] for c in pick]))
        for j, c in enumerate(pick):
            assert np.array_equal(got[j], ctx.answer(db, keys[c], cl[c]["query"])), j
        want = oracle.answer(r=r, n1=n1, db=db_words, query=cl[1]["query"], rgsw=cl[1]["rgsw"], baby=cl[1]["baby"],
                             giant=cl[1]["giant"], packing=cl[1]["packing"], perms=perms)
        assert np.array_equal(got[1], want)
        # argument errors come back with their reason
        with pytest.raises(RuntimeError, match="max_batch"):   # KspirError
            ctx.autotune(db, keys, max_batch=0)
    finally:
        for k, v in (("mfma_min_batch", 2), ("mfma_max_chunk", 1 << 30), ("fused_min_batch", 16),
                     ("fused_pack", 1), ("db_waves", 4), ("giant_wgs_per_cu", 1), ("giant_dma", 1),
                     ("pack_fused_min_wgs", 0)):
            ctx.set(k, v)
        for k in keys:
            k.close()
        db.close()


@pytest.mark.parametrize("r,n1,seeds", [(4, 64, (21, 22, 23)), (16, 128, (5,)), (1, 256, (9,))])
def test_client_generation_matches_oracle(ctx, oracle, r, n1, seeds):
    """secret, query, RGSW, baby/giant and packing keys generated on the GPU
    (rand() stream by jump-ahead) equal the oracle's client generation word
    for word; the GPU-made client's answer equals the oracle's"""
    say(f"[gpu] client generation r={r} n1={n1} clients={len(seeds)}")
    tcols = [(97 * i + 13) % (N // 2) for i in range(len(seeds))]
    tpacks = [(5 * i + 1) % r for i in range(len(seeds))]
    got = ctx.client_generate(r, n1, seeds, tcols, tpacks)
    for i, sd in enumerate(seeds):
        want = oracle.gen_client(r, n1, sd, tcols[i], tpacks[i])
        for k in ("secret", "query", "rgsw", "baby", "giant", "packing"):
            assert np.array_equal(got[k][i].reshape(want[k].shape), want[k]), (sd, k)
    if r == 4:
        db_words = oracle.gen_db(r, n1, seed=77)
        perms = oracle.perm_matrix(n1)
        db = ctx.database(db_words, r, n1)
        keys = ctx.keys(r, n1, got["baby"][0], got["giant"][0], got["packing"][0], got["rgsw"][0], perms)
        resp = ctx.answer(db, keys, got["query"][0])
        want = oracle.answer(r=r, n1=n1, db=db_words, query=got["query"][0], rgsw=got["rgsw"][0],
                             baby=got["baby"][0], giant=got["giant"][0], packing=got["packing"][0], perms=perms)
        assert np.array_equal(resp, want)
        keys.close()
        db.close()


@pytest.mark.parametrize("r,nq", [(64, 16), (256, 2), (256, 8), (1024, 2)])
def test_full_size_round_trip(ctx, oracle, r, nq):
    """1 GB (r 64), 4 GB (r 256) and 16 GB (r 1024) databases preprocessed on the GPU from one
    planted record block repeated r times, clients generated on the GPU,
    answered one by one (VALU database pass) and as one batch (int8-MFMA pass;
    at r 64 with 16 queries, r 256 with 8 and r 1024 with 2 the chunked
    pipeline with the fused giant-step and packing kernels, which a large
    database takes from a few queries on), recovered on the GPU: the batch
    equals the single answers word for word and every client recovers its
    planted message"""
    from tests.planted import planted_block, planted_message
    n1 = 128
    say(f"[gpu] full-size round trip r={r} ({r * 16} MB), {nq} clients")
    cols = [(97 * c + 5) % (N // 2) for c in range(nq)]
    packs = [(13 * c + r - 1) % r for c in range(nq)]
    shifts = [7 * c for c in range(nq)]
    rec = planted_block(oracle, r, cols, packs, shifts)
    db = ctx.database_from_records(rec[None], r, n1, block_stride=0)
    perms = oracle.perm_matrix(n1)
    cl = ctx.client_generate(r, n1, [300 + c for c in range(nq)], cols, packs)
    keys = [ctx.keys(r, n1, cl["baby"][c], cl["giant"][c], cl["packing"][c], cl["rgsw"][c], perms) for c in range(nq)]
    try:
        single = np.stack([ctx.answer(db, keys[c], cl["query"][c]) for c in range(nq)])
        if nq > 1:
            batch = ctx.answer_batch(db, keys, cl["query"])
            assert np.array_equal(batch, single)
        msgs = ctx.decrypt_bsgs_total(r, cl["secret"], single)
        for c in range(nq):
            assert np.array_equal(msgs[c], planted_message(r, shifts[c])), c
    finally:
        for k in keys:
            k.close()
        db.close()


def test_cpp_client_dropin_matches_oracle(oracle, tmp_path):
    """the client side of test-pir through kspir_pir.h (srand, Secret,
    query_bsgs_rns, RGSW / BSGS / packing keyGen, rand() drawn on the host in
    the reference's order) gives the oracle's words; decrypt_bsgs_total
    recovers the planted record and leaves the response in NTT form"""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "native", "_build", "dropin_test")
    if not os.path.exists(exe):
        pytest.fail("tests/native/_build/dropin_test not built (python kspir_amd/build.py)")
    r, n1, seed, tcol, tpack = 4, 64, 41, 300, 2
    say(f"[gpu] C++ client drop-in r={r} n1={n1} seed={seed}")
    out = subprocess.run([exe, "client", str(tmp_path), str(r), str(n1), str(seed), str(tcol), str(tpack)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    want = oracle.gen_client(r, n1, seed, tcol, tpack)
    for k in ("secret", "query", "rgsw", "baby", "giant", "packing"):
        got = np.fromfile(str(tmp_path / f"{k}.bin"), dtype=np.uint64).reshape(want[k].shape)
        assert np.array_equal(got, want[k]), k
    inst, resp, _ = instance(oracle, 2, 128, 3)     # gen_instance(seed 3): its secret is srand(3)'s
    resp.tofile(str(tmp_path / "response.bin"))
    out = subprocess.run([exe, "decrypt", str(tmp_path), "2", "3"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    msg = np.fromfile(str(tmp_path / "message.bin"), dtype=np.uint64)
    assert np.array_equal(msg, np.arange(1, N + 1, dtype=np.uint64))
    after = np.fromfile(str(tmp_path / "response_after.bin"), dtype=np.uint64).reshape(2, N)
    assert np.array_equal(after[0], oracle.ntt_forward(0, resp[0]))
    assert np.array_equal(after[1], oracle.ntt_forward(0, resp[1]))


def test_cpp_dropin_matches_oracle(oracle, tmp_path):
    """tests/native/dropin_test calls matrix_vector_mul_bsgs_rns_crt_large,
    externalProduct and packingRLWEs through kspir_pir.h exactly as
    test-pir.cpp:111-115 does; its response words, first-dimension results
    and mutated query equal the oracle's"""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "native", "_build", "dropin_test")
    if not os.path.exists(exe):
        pytest.fail("tests/native/_build/dropin_test not built (python kspir_amd/build.py)")
    r, n1 = 4, 128
    say(f"[gpu] C++ drop-in r={r} n1={n1}")
    inst, resp, st = instance(oracle, r, n1, 21)
    for name in ("db", "query", "rgsw", "baby", "giant", "packing", "perms"):
        getattr(inst, name).tofile(str(tmp_path / f"{name}.bin"))
    rec, rec_words = oracle.gen_records(r, n1, seed=33, plant=True, tcol=7, tpack=1)
    rec.tofile(str(tmp_path / "records.bin"))
    out = subprocess.run([exe, str(tmp_path), str(r), str(n1)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    say(out.stdout.strip())
    got = np.fromfile(str(tmp_path / "response.bin"), dtype=np.uint64).reshape(2, N)
    fd = np.fromfile(str(tmp_path / "firstdim.bin"), dtype=np.uint64).reshape(r, 2, N)
    q0 = np.fromfile(str(tmp_path / "query0_after.bin"), dtype=np.uint64).reshape(2, N)
    assert np.array_equal(fd, st["firstdim"])
    assert np.array_equal(got, resp)
    cb = st["cipherbuf"][:, 0, :]
    lo, hi = (cb & np.uint64(0xffffffff)), (cb >> np.uint64(32))
    want_q0 = np.array([[int(oracle.lib().ko_crt_compose(int(lo[z, ab]), int(hi[z, ab]))) for z in range(N)]
                        for ab in range(2)], dtype=np.uint64)
    assert np.array_equal(q0, want_q0)
    assert "input[0] ntt=1 result ntt=0" in out.stdout
    msg = oracle.decrypt(r, inst.secret, got)
    assert np.array_equal(msg, np.arange(1, N + 1, dtype=np.uint64))
    # database_tobsgsntt + database_tocrt through the drop-in give the oracle's words
    pre = np.fromfile(str(tmp_path / "datacrt_gpu.bin"), dtype=np.uint64).reshape(r, N * N // 2)
    assert np.array_equal(pre, rec_words)
    # a database word edited in place between two calls: the second answer
    # follows the edited words (the oracle's first dimension on them)
    w, val = np.fromfile(str(tmp_path / "edit_index.bin"), dtype=np.uint64)
    edited = inst.db.copy()
    assert edited.reshape(-1)[int(w)] != val
    edited.reshape(-1)[int(w)] = val
    _, st2 = oracle.answer(r=r, n1=n1, db=edited, query=inst.query, rgsw=inst.rgsw, baby=inst.baby,
                           giant=inst.giant, packing=inst.packing, perms=inst.perms, want_stages=True)
    fd2 = np.fromfile(str(tmp_path / "firstdim_edit.bin"), dtype=np.uint64).reshape(r, 2, N)
    assert not np.array_equal(st2["firstdim"], st["firstdim"])
    assert np.array_equal(fd2, st2["firstdim"])


def test_cpp_sample_database_matches_oracle(oracle, tmp_path):
    """sample_database_bsgs through kspir_pir.h after srand(seed) draws the
    oracle's records (test-pir.cpp:58 order, twosteps.cpp:23-39)"""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "native", "_build", "dropin_test")
    if not os.path.exists(exe):
        pytest.fail("tests/native/_build/dropin_test not built (python kspir_amd/build.py)")
    say("[gpu] C++ sample_database_bsgs r=2 seed=19")
    out = subprocess.run([exe, "sample", str(tmp_path), "2", "19"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = np.fromfile(str(tmp_path / "records.bin"), dtype=np.uint64).reshape(2, N, N // 2)
    want, _ = oracle.gen_records(2, 128, seed=19, plant=False)
    assert np.array_equal(got, want)


def test_reference_test_pir_unchanged(tmp_path):
    """the reference's own tests/test-pir.cpp, compiled unchanged against
    the drop-in headers (kspir_amd/cpp/compat) and linked with
    libkspir_amd.so, at its default 256 MB shape: it draws the README's
    targets and recovers the planted record [1..4096] (README.md:56,65)"""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "native", "_build", "test_pir")
    if not os.path.exists(exe):
        pytest.skip("the reference's tests/test-pir.cpp was not present when the library was built")
    say("[gpu] reference test-pir driver on the drop-in")
    out = subprocess.run([exe], cwd=str(tmp_path), capture_output=True, text=True, timeout=600)
    say(out.stdout.strip())
    assert out.returncode == 0, out.stderr
    assert "target_col: 177, target_packing: 4" in out.stdout
    assert ("the recovered result = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, ..., 4096]"
            in out.stdout)


def test_inverse_encode_on_gpu(ctx, oracle):
    """inverse_encode (forward NTT mod p, query_decode order) on the GPU
    equals the oracle's, including all-zero and all-(p-1) columns; values
    >= p are rejected"""
    import kspir_amd
    say("[gpu] inverse_encode")
    rng = np.random.default_rng(23)
    cols = rng.integers(0, 65537, (4, N), dtype=np.uint64)
    cols[0] = 0
    cols[1] = 65536
    got = ctx.inverse_encode(cols)
    for i in range(4):
        assert np.array_equal(got[i], oracle.inverse_encode(cols[i])), i
    bad = cols[2].copy()
    bad[7] = 65537
    with pytest.raises(kspir_amd.KspirError, match="below p"):
        ctx.inverse_encode(bad)


def test_database_compact_keeps_answers_and_export(ctx, oracle):
    """a database kept as int8 limbs only (kspir_database_compact) gives the
    same single and batched answers, exports the same words, and is rejected
    by the entry points that need its words"""
    import kspir_amd
    r, n1 = 4, 128
    say(f"[gpu] compacted database r={r} n1={n1}")
    words = oracle.gen_db(r, n1, seed=21)
    db = ctx.database(words, r, n1)
    perms = oracle.perm_matrix(n1)
    cls = [oracle.gen_client(r, n1, 50 + c, (31 * c + 3) % (N // 2), c % r) for c in range(3)]
    keys = [ctx.keys(r, n1, cl["baby"], cl["giant"], cl["packing"], cl["rgsw"], perms) for cl in cls]
    try:
        qs = np.stack([cl["query"] for cl in cls])
        single = [ctx.answer(db, keys[c], qs[c]) for c in range(3)]
        batch = ctx.answer_batch(db, keys, qs)
        db.compact()
        db.compact()   # idempotent
        assert np.array_equal(db.export(), words)
        for c in range(3):
            assert np.array_equal(ctx.answer(db, keys[c], qs[c]), single[c]), c
        assert np.array_equal(ctx.answer_batch(db, keys, qs), batch)
        assert np.array_equal(single[0], oracle.answer(r=r, n1=n1, db=words, query=qs[0], rgsw=cls[0]["rgsw"],
                                                       baby=cls[0]["baby"], giant=cls[0]["giant"],
                                                       packing=cls[0]["packing"], perms=perms))
        qb, rb = ctx.alloc(4 * N * 8), ctx.alloc(2 * N * 8)
        try:
            qb.upload(qs[0])
            with pytest.raises(kspir_amd.KspirError, match="limbs only"):
                ctx.answer_device_noalloc(db, keys[0], qb, rb)
        finally:
            qb.free()
            rb.free()
    finally:
        for k in keys:
            k.close()
        db.close()


def test_database_file_round_trip(ctx, oracle, tmp_path):
    """kspir_database_save / _load: the reloaded database answers and exports
    the same words; a file with a wrong magic, a truncated file or a flipped
    word is rejected"""
    import kspir_amd
    r, n1 = 2, 128
    say(f"[gpu] database file round trip r={r} n1={n1}")
    words = oracle.gen_db(r, n1, seed=33)
    db = ctx.database(words, r, n1)
    path = str(tmp_path / "db.kspir")
    db.save(path)
    db2 = ctx.database_load(path)
    cl = oracle.gen_client(r, n1, 71, 12, 1)
    keys = ctx.keys(r, n1, cl["baby"], cl["giant"], cl["packing"], cl["rgsw"], oracle.perm_matrix(n1))
    try:
        assert (db2.r, db2.n1) == (r, n1)
        assert np.array_equal(db2.export(), words)
        assert np.array_equal(ctx.answer(db2, keys, cl["query"]), ctx.answer(db, keys, cl["query"]))
        raw = open(path, "rb").read()
        bad = [b"XSPIRDB\0" + raw[8:], raw[:-8], raw[:64 + 800] + bytes([raw[64 + 800] ^ 1]) + raw[64 + 801:]]
        for i, blob in enumerate(bad):
            p2 = str(tmp_path / f"bad{i}.kspir")
            open(p2, "wb").write(blob)
            with pytest.raises(kspir_amd.KspirError):
                ctx.database_load(p2)
    finally:
        keys.close()
        db2.close()
        db.close()
