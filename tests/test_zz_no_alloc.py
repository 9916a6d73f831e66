"""kspir_answer_device_noalloc (the entry point that allocates no device
memory) against kspir_answer_device on the same device buffers: word-equal
responses on consecutive calls, the second with a smaller database shape, and
a shape above the fixed scratch rejected with an error before any launch."""
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


def _setup(ctx, oracle, r, n1, seeds):
    """planted database of shape (r, n1) preprocessed on the GPU and clients
    generated on the GPU for it"""
    nq = len(seeds)
    cols = [(53 * c + 11) % (N // 2) for c in range(nq)]
    packs = [(7 * c + r - 1) % r for c in range(nq)]
    shifts = [5 * c + 2 for c in range(nq)]
    rec = planted_block(oracle, r, cols, packs, shifts)
    db = ctx.database_from_records(rec[None], r, n1, block_stride=0)
    perms = oracle.perm_matrix(n1)
    cl = ctx.client_generate(r, n1, list(seeds), cols, packs)
    keys = [ctx.keys(r, n1, cl["baby"][c], cl["giant"][c], cl["packing"][c], cl["rgsw"][c], perms)
            for c in range(nq)]
    return db, keys, cl, shifts


def test_noalloc_matches_answer_device(ctx, oracle):
    qb, rb_ref, rb_na = ctx.alloc(4 * N * 8), ctx.alloc(2 * N * 8), ctx.alloc(2 * N * 8)
    handles = []
    try:
        for r, n1, seeds in ((16, 128, (401, 402)), (4, 64, (403,))):
            say(f"[gpu] no-alloc answer r={r} n1={n1} clients={len(seeds)}")
            db, keys, cl, shifts = _setup(ctx, oracle, r, n1, seeds)
            handles.append((db, keys))
            for c in range(len(seeds)):
                qb.upload(cl["query"][c])
                ctx.answer_device(db, keys[c], qb, rb_ref)
                ctx.synchronize()
                want = rb_ref.download(np.zeros((2, N), dtype=np.uint64))
                rb_na.upload(np.zeros((2, N), dtype=np.uint64))
                ctx.answer_device_noalloc(db, keys[c], qb, rb_na)
                ctx.synchronize()
                got = rb_na.download(np.zeros((2, N), dtype=np.uint64))
                assert np.array_equal(got, want), (r, n1, c)
                msg = ctx.decrypt_bsgs_total(r, cl["secret"][c:c + 1], got[None])[0]
                assert np.array_equal(msg, planted_message(r, shifts[c])), (r, n1, c)
        # back to the larger shape after the smaller one: stale scratch
        # contents of the previous call do not leak into the response
        db, keys = handles[0]
        say("[gpu] no-alloc answer r=16 n1=128 again")
        cl = ctx.client_generate(16, 128, [401], [11], [15])
        k2 = ctx.keys(16, 128, cl["baby"][0], cl["giant"][0], cl["packing"][0], cl["rgsw"][0],
                      oracle.perm_matrix(128))
        handles.append((None, [k2]))
        qb.upload(cl["query"][0])
        ctx.answer_device(db, k2, qb, rb_ref)
        ctx.answer_device_noalloc(db, k2, qb, rb_na)
        ctx.synchronize()
        assert np.array_equal(rb_na.download(np.zeros((2, N), dtype=np.uint64)),
                              rb_ref.download(np.zeros((2, N), dtype=np.uint64)))
    finally:
        for db, keys in handles:
            for k in keys:
                k.close()
            if db is not None:
                db.close()
        for b in (qb, rb_ref, rb_na):
            b.free()


def test_noalloc_rejects_shapes_above_scratch(ctx, oracle):
    """r 16 at N1 64 needs about 92 MB of workspace: rejected on the host"""
    import kspir_amd
    say("[gpu] no-alloc shape check r=16 n1=64")
    r, n1 = 16, 64
    n2 = N // 2 // n1
    db = ctx.database(np.zeros((r, N * N // 2), dtype=np.uint64), r, n1)
    z = np.zeros
    keys = ctx.keys(r, n1, z((n1 // 2, 6, 2, N), np.uint64), z((n2 - 1, 2, 2, N), np.uint64),
                    z((4, 4, 2, N), np.uint64), z((4, 2, N), np.uint64), oracle.perm_matrix(n1))
    qb, rb = ctx.alloc(4 * N * 8), ctx.alloc(2 * N * 8)
    try:
        with pytest.raises(kspir_amd.KspirError, match="scratch"):
            ctx.answer_device_noalloc(db, keys, qb, rb)
    finally:
        keys.close()
        db.close()
        qb.free()
        rb.free()
