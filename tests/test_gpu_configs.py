"""GPU parity at the benchmark configurations' own shapes.

  * 64 clients with distinct keys in one batch against the 256 MB database
    (r 16, N1 128): the configuration the headline throughput is timed on,
    with the default options (one int8-MFMA sweep of 128 columns, the fused
    giant-step and packing kernels, hipGraph replay);
  * three databases of uneven size, 64 MB / 512 MB / 4 GB (r 4 / 32 / 256),
    with 16 / 32 / 48 clients, batched (int8-MFMA pass) against the clients'
    single answers (VALU pass), one after the other and overlapped on three
    contexts;
  * the serving loop on the 4 GB database with a burst of 64 submissions.

Every database is preprocessed on the GPU from a planted record block
(tests/planted.py), every client is generated on the GPU, and every response
must (a) equal the same query's single answer word for word, (b) decrypt to
the client's planted message; for the 256 MB and 64 MB shapes two responses
must also equal the oracle's.  Exact integer arithmetic: tolerance 0."""
import sys

import numpy as np
import pytest

from tests.planted import planted_block, planted_message

pytestmark = pytest.mark.gpu

N = 4096


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import kspir_amd
    c = kspir_amd.Context(0)
    yield c
    c.close()


class Shape:
    """a planted database of shape (r, N1) resident on the GPU and nq clients"""

    def __init__(self, ctx, oracle, r, n1, nq, seed0):
        self.r, self.n1, self.nq = r, n1, nq
        self.cols = [(97 * c + 5 + seed0) % (N // 2) for c in range(nq)]
        self.packs = [(13 * c + r - 1) % r for c in range(nq)]
        self.shifts = [7 * c + 1 for c in range(nq)]
        rec = planted_block(oracle, r, self.cols, self.packs, self.shifts)
        self.db = ctx.database_from_records(rec[None], r, n1, block_stride=0)
        self.perms = oracle.perm_matrix(n1)
        self.cl = ctx.client_generate(r, n1, [seed0 + c for c in range(nq)], self.cols, self.packs)
        self.keys = [ctx.keys(r, n1, self.cl["baby"][c], self.cl["giant"][c], self.cl["packing"][c],
                              self.cl["rgsw"][c], self.perms) for c in range(nq)]
        self.queries = self.cl["query"]
        self.single = np.stack([ctx.answer(self.db, self.keys[c], self.queries[c]) for c in range(nq)])

    def check_decrypts(self, ctx, resp):
        msgs = ctx.decrypt_bsgs_total(self.r, self.cl["secret"], resp)
        for c in range(self.nq):
            assert np.array_equal(msgs[c], planted_message(self.r, self.shifts[c])), c

    def check_oracle(self, oracle, resp, picks):
        words = self.db.export()
        for c in picks:
            want = oracle.answer(r=self.r, n1=self.n1, db=words, query=self.queries[c], rgsw=self.cl["rgsw"][c],
                                 baby=self.cl["baby"][c], giant=self.cl["giant"][c],
                                 packing=self.cl["packing"][c], perms=self.perms)
            assert np.array_equal(resp[c], want), c

    def close(self):
        for k in self.keys:
            k.close()
        self.db.close()


@pytest.fixture(scope="module")
def db4gb(ctx, oracle):
    say("[gpu] 4 GB database (r=256, n1=128) and 48 clients")
    s = Shape(ctx, oracle, 256, 128, 48, 3000)
    yield s
    s.close()


def test_batch_64_clients_256MB(ctx, oracle):
    """the headline configuration: 64 queries from 64 clients in one batch"""
    say("[gpu] 256 MB (r=16, n1=128), 64 clients in one batch")
    s = Shape(ctx, oracle, 16, 128, 64, 1)
    try:
        assert ctx.get("mfma_max_chunk") >= 64   # one database sweep for the whole batch
        batch = ctx.answer_batch(s.db, s.keys, s.queries)
        for c in range(s.nq):
            assert np.array_equal(batch[c], s.single[c]), c
        s.check_decrypts(ctx, batch)
        # the device-buffer entry point, eager then two graph replays
        qb, rb = ctx.alloc(s.nq * 4 * N * 8), ctx.alloc(s.nq * 2 * N * 8)
        try:
            qb.upload(s.queries)
            for rep in range(3):
                rb.upload(np.zeros((s.nq, 2, N), dtype=np.uint64))
                ctx.answer_batch_device(s.db, s.keys, qb, rb)
                ctx.synchronize()
                assert np.array_equal(rb.download(np.zeros((s.nq, 2, N), dtype=np.uint64)), batch), rep
        finally:
            qb.free()
            rb.free()
        say("[gpu] 256 MB batch: oracle on clients 0 and 63")
        s.check_oracle(oracle, batch, (0, 63))
    finally:
        s.close()


def test_uneven_databases_batched(ctx, oracle, db4gb):
    """64 MB / 512 MB / 4 GB with 16 / 32 / 48 clients: each database's batch
    (int8-MFMA pass) equals its clients' single answers (VALU pass) and
    decrypts; the three batches overlapped on three contexts give the same
    words"""
    import kspir_amd
    say("[gpu] 64 MB (r=4) and 512 MB (r=32) databases, 16 and 32 clients")
    small = [Shape(ctx, oracle, 4, 128, 16, 1000), Shape(ctx, oracle, 32, 128, 32, 2000)]
    shapes = small + [db4gb]
    try:
        batches = []
        for s in shapes:
            say(f"[gpu] batch r={s.r} nq={s.nq}")
            b = ctx.answer_batch(s.db, s.keys, s.queries)
            for c in range(s.nq):
                assert np.array_equal(b[c], s.single[c]), (s.r, c)
            s.check_decrypts(ctx, b)
            batches.append(b)
        small[0].check_oracle(oracle, batches[0], (0, 15))
        # one batch spanning the three databases: one launch per stage
        say("[gpu] one batch spanning the three databases")
        segs = [(s.db, s.keys) for s in shapes]
        allq = np.concatenate([s.queries for s in shapes])
        multi = ctx.answer_multi(segs, allq)
        assert np.array_equal(multi, np.concatenate(batches))
        # the device-buffer entry point, eager then a graph replay
        qb, rb = ctx.alloc(len(allq) * 4 * N * 8), ctx.alloc(len(allq) * 2 * N * 8)
        try:
            qb.upload(allq)
            for rep in range(2):
                rb.upload(np.zeros((len(allq), 2, N), dtype=np.uint64))
                ctx.answer_multi_device(segs, qb, rb)
                ctx.synchronize()
                assert np.array_equal(rb.download(np.zeros((len(allq), 2, N), dtype=np.uint64)), multi), rep
        finally:
            qb.free()
            rb.free()
        # overlapped: one context (stream) per database, all enqueued before any wait
        say("[gpu] three batches overlapped on three contexts")
        others = [kspir_amd.Context(0), kspir_amd.Context(0)]
        ctxs = [ctx] + others
        bufs = []
        try:
            for s in shapes:
                qb, rb = ctx.alloc(s.nq * 4 * N * 8), ctx.alloc(s.nq * 2 * N * 8)
                qb.upload(s.queries)
                rb.upload(np.zeros((s.nq, 2, N), dtype=np.uint64))
                bufs.append((qb, rb))
            for c, s, (qb, rb) in zip(ctxs, shapes, bufs):
                c.answer_batch_device(s.db, s.keys, qb, rb)
            for c in ctxs:
                c.synchronize()
            for s, (qb, rb), b in zip(shapes, bufs, batches):
                assert np.array_equal(rb.download(np.zeros((s.nq, 2, N), dtype=np.uint64)), b), s.r
        finally:
            for qb, rb in bufs:
                qb.free()
                rb.free()
            for c in others:
                c.close()
    finally:
        for s in small:
            s.close()


def test_serving_burst_4GB(ctx, oracle, db4gb):
    """64 submissions in a burst to the serving loop over the 4 GB database
    (48 clients, 16 of them twice): every response equals the client's single
    answer; admission by footprint splits the burst into batches"""
    import kspir_amd
    s = db4gb
    say("[gpu] serving loop, 4 GB database, burst of 64")
    order = list(range(s.nq)) + list(range(16))
    # a budget that admits about 24 queries of this shape per batch
    srv = kspir_amd.Server(ctx, max_batch=64, budget_bytes=24 * kspir_amd.query_footprint(s.r, s.n1))
    try:
        tickets = [(c, srv.submit(s.db, s.keys[c], s.queries[c])) for c in order]
        lat = []
        for c, t in tickets:
            resp, ms = srv.wait(t)
            assert np.array_equal(resp, s.single[c]), c
            lat.append(ms)
        st = srv.stats()
        assert st["answered"] == len(order) and st["pending"] == 0 and st["batches"] >= 3
        say(f"[gpu] serving burst: {st['batches']} batches, p50 {np.percentile(lat, 50):.1f} ms, "
            f"max {max(lat):.1f} ms")
    finally:
        srv.close()


@pytest.mark.parametrize("chunk", [1, 2, 4, 0])
def test_chunked_batch_pipeline(ctx, oracle, chunk):
    """the batched pipeline run a chunk of blocks at a time (chunk_blocks 1,
    2, 4; 0 = the whole database at once) gives the single answers' words
    for 16 clients on an r 8 database, and the oracle's for client 0"""
    say(f"[gpu] chunked batch r=8 n1=128 nq=16 chunk_blocks={chunk}")
    s = Shape(ctx, oracle, 8, 128, 16, 4000)
    ctx.set("chunk_blocks", chunk)
    try:
        batch = ctx.answer_batch(s.db, s.keys, s.queries)
    finally:
        ctx.set("chunk_blocks", 32)
    try:
        for c in range(s.nq):
            assert np.array_equal(batch[c], s.single[c]), c
        s.check_decrypts(ctx, batch)
        if chunk == 1:
            s.check_oracle(oracle, batch, (0,))
    finally:
        s.close()


def test_serving_rejects_double_wait(ctx, oracle):
    """a ticket is collected once: a second wait on it, or a wait on a ticket
    never issued, is an error and not a hang"""
    import kspir_amd
    say("[gpu] serving loop double wait")
    r, n1 = 2, 128
    db = ctx.database(oracle.gen_db(r, n1, seed=5), r, n1)
    cl = oracle.gen_client(r, n1, 77, 9, 1)
    keys = ctx.keys(r, n1, cl["baby"], cl["giant"], cl["packing"], cl["rgsw"], oracle.perm_matrix(n1))
    want = ctx.answer(db, keys, cl["query"])
    srv = kspir_amd.Server(ctx, max_batch=4)
    try:
        t = srv.submit(db, keys, cl["query"])
        resp, _ = srv.wait(t)
        assert np.array_equal(resp, want)
        srv._live[t] = (None, resp)   # the binding keeps no record of collected tickets
        with pytest.raises(kspir_amd.KspirError, match="ticket"):
            srv.wait(t)
        with pytest.raises(kspir_amd.KspirError, match="ticket"):
            srv.wait(t + 1000)
    finally:
        srv.close()
        keys.close()
        db.close()


def test_serving_failure_paths(ctx, oracle):
    """failures forced from the host, every kernel exiting normally: a query
    no batch could hold is rejected at submission (EBUDGET); requests that
    no batch has taken yet are withdrawn (their waits raise ECANCELED, the
    others answer correctly), a cancel too late or of an unknown ticket is an
    error; a batch failed on purpose (serve_fail_batches) fails only its own
    requests, and the loop answers the next ones"""
    import kspir_amd
    say("[gpu] serving loop failure paths")
    r, n1 = 2, 128
    db = ctx.database(oracle.gen_db(r, n1, seed=6), r, n1)
    cl = oracle.gen_client(r, n1, 78, 11, 0)
    keys = ctx.keys(r, n1, cl["baby"], cl["giant"], cl["packing"], cl["rgsw"], oracle.perm_matrix(n1))
    want = ctx.answer(db, keys, cl["query"])
    one = kspir_amd.batch_footprint(r, n1, 1, ctx)
    try:
        # admission: a budget below one query's workspace
        srv = kspir_amd.Server(ctx, max_batch=4, budget_bytes=one // 2)
        try:
            with pytest.raises(kspir_amd.KspirError) as e:
                srv.submit(db, keys, cl["query"])
            assert e.value.code == kspir_amd.EBUDGET
        finally:
            srv.close()
        # cancellation: one query per batch, so the late submissions queue
        srv = kspir_amd.Server(ctx, max_batch=1)
        try:
            tickets = [srv.submit(db, keys, cl["query"]) for _ in range(16)]
            cancelled = []
            for t in tickets[-6:]:
                try:
                    srv.cancel(t)
                    cancelled.append(t)
                except kspir_amd.KspirError as e:
                    assert e.code == kspir_amd.EBUSY   # already taken by a batch
            assert len(cancelled) >= 2
            for t in tickets:
                if t in cancelled:
                    with pytest.raises(kspir_amd.KspirError) as e:
                        srv.wait(t)
                    assert e.value.code == kspir_amd.ECANCELED
                else:
                    resp, _ = srv.wait(t)
                    assert np.array_equal(resp, want)
            st = srv.stats()
            assert st["answered"] == 16 - len(cancelled) and st["pending"] == 0
            # collected, or never issued
            with pytest.raises(kspir_amd.KspirError) as e:
                srv.cancel(tickets[0])
            assert e.value.code == kspir_amd.EINVAL
            with pytest.raises(kspir_amd.KspirError) as e:
                srv.cancel(tickets[-1] + 1000)
            assert e.value.code == kspir_amd.EINVAL
        finally:
            srv.close()
        # a failed batch: its own requests fail, the next ones are answered
        ctx.set("serve_fail_batches", 1)
        srv = kspir_amd.Server(ctx, max_batch=1)
        try:
            with pytest.raises(kspir_amd.KspirError) as e:
                srv.wait(srv.submit(db, keys, cl["query"]))
            assert e.value.code == kspir_amd.EINJECTED
            resp, _ = srv.wait(srv.submit(db, keys, cl["query"]))
            assert np.array_equal(resp, want)
        finally:
            srv.close()
        assert ctx.get("serve_fail_batches") == 0
    finally:
        ctx.set("serve_fail_batches", 0)
        keys.close()
        db.close()


@pytest.mark.parametrize("chunk,sweep", [(32, 1 << 30), (1, 4), (2, 1 << 30)])
def test_multi_database_batch(ctx, oracle, chunk, sweep):
    """one batch spanning databases of r 1, 4 and 2 (in that order) with 3, 5
    and 2 clients: every response equals its single answer against its own
    database and decrypts, one per database equals the oracle's; with
    chunk_blocks 1 and 2 and with MFMA sweeps of 4 queries (databases split
    across sweeps); device buffers replayed through the captured graph"""
    say(f"[gpu] batch spanning r=1/4/2 databases, chunk_blocks={chunk} mfma_max_chunk={sweep}")
    shapes = [Shape(ctx, oracle, 1, 128, 3, 5000), Shape(ctx, oracle, 4, 128, 5, 6000),
              Shape(ctx, oracle, 2, 128, 2, 7000)]
    ctx.set("chunk_blocks", chunk)
    ctx.set("mfma_max_chunk", sweep)
    try:
        segs = [(s.db, s.keys) for s in shapes]
        allq = np.concatenate([s.queries for s in shapes])
        multi = ctx.answer_multi(segs, allq)
        qb, rb = ctx.alloc(len(allq) * 4 * N * 8), ctx.alloc(len(allq) * 2 * N * 8)
        try:
            qb.upload(allq)
            for rep in range(3):
                rb.upload(np.zeros((len(allq), 2, N), dtype=np.uint64))
                ctx.answer_multi_device(segs, qb, rb)
                ctx.synchronize()
                assert np.array_equal(rb.download(np.zeros((len(allq), 2, N), dtype=np.uint64)), multi), rep
        finally:
            qb.free()
            rb.free()
    finally:
        ctx.set("chunk_blocks", 32)
        ctx.set("mfma_max_chunk", 1 << 30)
    try:
        q = 0
        for s in shapes:
            resp = multi[q:q + s.nq]
            for c in range(s.nq):
                assert np.array_equal(resp[c], s.single[c]), (s.r, c)
            s.check_decrypts(ctx, resp)
            if chunk == 32:
                s.check_oracle(oracle, resp, (s.nq - 1,))
            q += s.nq
    finally:
        for s in shapes:
            s.close()


def test_multi_database_rejects_mixed_n1(ctx, oracle):
    """the databases of one batch share N1; a mismatch is an error"""
    import kspir_amd
    a = Shape(ctx, oracle, 2, 128, 1, 8000)
    b = Shape(ctx, oracle, 2, 64, 1, 8100)
    try:
        with pytest.raises(kspir_amd.KspirError, match="N1"):
            ctx.answer_multi([(a.db, a.keys), (b.db, b.keys)], np.concatenate([a.queries, b.queries]))
    finally:
        a.close()
        b.close()
