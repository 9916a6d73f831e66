#!/usr/bin/env python3
"""Benchmark of the MI355X KsPIR online answer.

One step = one batched online server response: B queries (default 64) from
B clients with their own keys, each answered as test-pir.cpp:107-116 does
(BSGS first dimension over the r database blocks, r external products,
packing), against one 256 MB database (32768 records x 8 KB: N 4096, r 16,
N1 128, N2 16).  Database, keys and queries are resident in HBM before
timing.  The plain single-query answer is timed in the same run and reported
beside it.  Inputs are synthetic (uniform residues of the right moduli; the
answer's cost does not depend on their values).  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch 64] [--db-mb 256] [--no-cpu-baseline]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 4096
Q1, Q2, BS = 268369921, 249561089, 16760833
Q = Q1 * Q2
HBM_PEAK_GBPS = 8000.0
README_SECONDS_256MB = 0.236635     # README.md:62 online response, 1 query, 256 MB (CPU, HEXL)


def synthetic_db(r, seed=1234):
    """uniform residue pairs in the database_tocrt word layout"""
    rng = np.random.default_rng(seed)
    words = r * N * N // 2
    lo = rng.integers(0, Q1, words, dtype=np.uint64)
    hi = rng.integers(0, Q2, words, dtype=np.uint64)
    return (lo | (hi << np.uint64(32))).reshape(r, -1)


def synthetic_records(seed=1234):
    """one database block of records in the sample_database_bsgs layout:
    [N][N/2] values < 2^16 (the reference samples rand() & 0xffff)"""
    rng = np.random.default_rng(seed)
    return rng.integers(0, 1 << 16, (1, N, N // 2), dtype=np.uint64)


def synthetic_client(r, n1, seed):
    """one client's query and keys: uniform residues of the right moduli in
    the reference layouts (see include/kspir_amd.h)"""
    rng = np.random.default_rng(seed)
    n2 = N // 2 // n1
    query = np.empty((2, 2, N), dtype=np.uint64)
    query[0] = rng.integers(0, Q, (2, N), dtype=np.uint64)
    query[1] = rng.integers(0, BS, (2, N), dtype=np.uint64)
    baby = np.empty((n1 // 2, 6, 2, N), dtype=np.uint64)
    baby[:, :3] = rng.integers(0, Q, (n1 // 2, 3, 2, N), dtype=np.uint64)
    baby[:, 3:] = rng.integers(0, BS, (n1 // 2, 3, 2, N), dtype=np.uint64)
    giant = rng.integers(0, Q, (n2 - 1, 2, 2, N), dtype=np.uint64)
    packing = rng.integers(0, Q, (max(r.bit_length() - 1, 0), 4, 2, N), dtype=np.uint64)
    rgsw = rng.integers(0, Q, (4, 2, N), dtype=np.uint64)
    return query, baby, giant, packing, rgsw


def traffic_from_profile(kernel="k_db_product"):
    """HBM bytes per launch of the DB-pass kernel from the committed rocprof
    PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)"""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def host_cpu():
    """CPU model and logical core count of this host"""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(r, n1, reps):
    """the CPU oracle (C++ restatement of the reference path) timed on this
    host: single thread, like test-pir, and all cores (one query per thread);
    inputs generated in test-pir order"""
    from oracle import oracle as O
    if not os.path.exists(O.LIB):
        O.build()
    model, ncpu = host_cpu()
    threads = max(1, min(16, ncpu))
    O.set_threads(threads)
    inst = O.gen_instance(r=r, n1=n1, seed=1)
    O.set_threads(1)
    first = O.answer(inst)   # warm-up
    times, stages = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.answer(inst)
        times.append(time.perf_counter() - t0)
        stages.append(O.last_stage_us())
    sec = float(np.median(times))
    names = ("baby_steps", "db_pass", "giant_steps", "external_products", "packing")
    split = {n: round(float(np.median([st[i] for st in stages])) / 1e3, 2) for i, n in enumerate(names)}
    loop = "AVX-512 database loop (crt.cpp:660-768 shape)" if O.db_product_vectorised() else "scalar database loop"
    loop += ", AVX-512 NTTs (8 lanes, as HEXL's AVX-512 path)" if O.ntt_set_vector(True) else ", scalar NTTs"
    single = {"value": round(1.0 / sec, 3), "unit": "queries/s", "cores": 1, "kind": "port",
              "sample": f"{reps} single-query online answers (median {sec * 1e3:.1f} ms) at r={r}, N1={n1} "
                        f"on oracle-generated test-pir inputs, seed 1, {loop}; host {model}, {ncpu} logical CPUs; "
                        f"the reference README reports 236,635 us per answer on unknown hardware (reported only)",
              "stages_ms": split}
    t0 = time.perf_counter()
    resp = O.answer_parallel(inst, threads, threads)
    wall = time.perf_counter() - t0
    assert np.array_equal(resp, first)
    allc = {"value": round(threads / wall, 3), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"{threads} answers of the same inputs, one per thread, {wall * 1e3:.0f} ms wall"}
    return single, allc


MFMA_I8_PEAK_TOPS = 5000.0     # MI355X dense int8 MFMA, 5.0 POP/s


def batched_roofline(r, n1, B, db3, direct=True):
    """the batched step's database pass (its kernels, timed live with HIP
    events) against its bounds: the MFMA kernel's int8 rate, and the pass's
    algorithmic HBM bytes over the kernels' time; VALU utilisation of the
    transform kernels from the committed PMC summary.  direct: the products
    are written in the first-dimension layout by k_db_mfma_direct (no move
    kernel); otherwise k_db_mfma + k_mfma_scatter"""
    if not db3:
        return None
    n2 = N // 2 // n1
    R, C = r * n2, (2 * B + 15) // 16 * 16
    ops = 2 * N * 2 * R * n1 * C * 16          # 2 primes x N slots x R x C outputs x n1 x 16 limb MACs x 2 ops
    limbs_ms, mfma_ms, move_ms = db3
    tops = ops / (mfma_ms * 1e-3) / 1e12
    # database limbs (8 B per word) + the batch's rotation words (4 B per
    # residue) in, the first-dimension products (4 B per residue) out
    alg = 8 * r * N * (N // 2) + B * n1 * 4 * N * 4 + B * R * 4 * N * 4
    pass_ms = limbs_ms + mfma_ms + move_ms
    name = "k_db_mfma_direct" if direct else "k_db_mfma"
    kernels = {"k_ct_limbs": round(limbs_ms, 4), name: round(mfma_ms, 4)}
    if not direct:
        kernels["k_mfma_scatter"] = round(move_ms, 4)
    out = {"kernel": name + " (batched int8-MFMA database pass)", "bound": "mfma",
           "achieved": round(tops, 1), "peak": MFMA_I8_PEAK_TOPS, "unit": "TOP/s (int8)",
           "frac": round(tops / MFMA_I8_PEAK_TOPS, 4), "ops_per_launch": ops, "avg_launch_ms": round(mfma_ms, 4),
           "db_pass": {"kernels_ms": kernels,
                       "algorithmic_bytes": alg, "achieved_GBps": round(alg / (pass_ms * 1e-3) / 1e9, 1),
                       "peak_GBps": HBM_PEAK_GBPS, "frac": round(alg / (pass_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)}}
    path = os.path.join(ROOT, "profiles", "batched_step_utilisation.json")
    if os.path.exists(path):
        with open(path) as f:
            out["transform_kernels_valu"] = json.load(f)
    return out


def config_c2(args, log):
    """C2: single query against a 1 GB database (r = 64, N1 = 128, N2 = 16),
    the database built on the GPU from one block of records repeated"""
    import kspir_amd
    r, n1 = 64, 128
    ctx = kspir_amd.Context(0)
    t0 = time.perf_counter()
    db = ctx.database_from_records(synthetic_records(), r, n1, block_stride=0)
    prep_s = time.perf_counter() - t0
    query, baby, giant, packing, rgsw = synthetic_client(r, n1, 7)
    keys = ctx.keys(r, n1, baby, giant, packing, rgsw, kspir_amd.compute_permutation_matrix(n1))
    qbuf, rbuf = ctx.alloc(4 * N * 8), ctx.alloc(2 * N * 8)
    qbuf.upload(query)
    ms = ctx.time_answer(db, keys, qbuf, rbuf, 5, args.plain_steps) / args.plain_steps
    lat = [ctx.time_answer(db, keys, qbuf, rbuf, 0, 1) for _ in range(args.lat_samples)]
    db_ms = ctx.time_db_pass(db, keys, qbuf, 10)
    db_bytes = 8 * r * N * (N // 2)
    out = {"config": "C2 single query, 1 GB database (131072 records x 8 KB), r=64, N1=128, N2=16",
           "qps": round(1000.0 / ms, 2), "ms": round(ms, 4),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                          "p99": round(float(np.percentile(lat, 99)), 4)},
           "db_pass": {"avg_launch_ms": round(db_ms, 4), "algorithmic_bytes": db_bytes,
                       "GBps": round(db_bytes / (db_ms * 1e-3) / 1e9, 1),
                       "frac_of_8TBps": round(db_bytes / (db_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)},
           "gpu_preprocess_s_incl_alloc": round(prep_s, 3)}
    keys.close()
    db.close()
    ctx.close()
    return out


def config_c4(args, log):
    """C4: uneven multi-database batch: 64 MB (r=4), 512 MB (r=32) and 4 GB
    (r=256) databases with 16, 32 and 48 queries (distinct clients), answered
    back to back; once with the int8-MFMA database pass, once with the VALU
    pass (both exact)"""
    import kspir_amd
    n1 = 128
    plan = [(4, 16), (32, 32), (256, 48)]
    ctx = kspir_amd.Context(0)
    perms = kspir_amd.compute_permutation_matrix(n1)
    rec = synthetic_records()
    groups, host_q = [], []
    seed = 5000
    for r, nq in plan:
        log(f"C4 database r={r}, {nq} clients")
        db = ctx.database_from_records(rec, r, n1, block_stride=0)
        keys, qs = [], []
        for _ in range(nq):
            query, baby, giant, packing, rgsw = synthetic_client(r, n1, seed)
            seed += 1
            keys.append(ctx.keys(r, n1, baby, giant, packing, rgsw, perms))
            qs.append(query)
        qbuf, rbuf = ctx.alloc(nq * 4 * N * 8), ctx.alloc(nq * 2 * N * 8)
        qbuf.upload(np.stack(qs))
        groups.append((r, nq, db, keys, qbuf, rbuf))
        host_q.append(np.stack(qs))
    total_q = sum(nq for _, nq in plan)
    out = {"config": "C4 uneven: 64 MB / 512 MB / 4 GB databases (r 4 / 32 / 256, N1 128) with 16 / 32 / 48 "
                     "queries of distinct clients, one step = all 96 answered back to back"}
    for mode, min_batch in (("mfma", 2), ("valu", 1 << 30)):
        ctx.set("mfma_min_batch", min_batch)
        per = {}
        for r, nq, db, keys, qbuf, rbuf in groups:       # warm-up (builds limbs, workspace)
            ctx.answer_batch_device(db, keys, qbuf, rbuf)
        ctx.synchronize()
        steps = 3
        t0 = time.perf_counter()
        for _ in range(steps):
            for r, nq, db, keys, qbuf, rbuf in groups:
                ctx.answer_batch_device(db, keys, qbuf, rbuf)
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / steps
        for r, nq, db, keys, qbuf, rbuf in groups:
            ms = ctx.time_answer_batch(db, keys, qbuf, rbuf, 1, 3) / 3
            per[f"r{r}_{nq}q"] = {"ms": round(ms, 3), "qps": round(nq * 1000.0 / ms, 1)}
        out[mode] = {"step_ms": round(wall * 1e3, 3), "qps": round(total_q / wall, 1), "per_database": per}
    ctx.set("mfma_min_batch", 2)
    # the same step with each database's batch on a context (stream) of its
    # own, all enqueued before any is waited for, so the small databases'
    # batches overlap the 4 GB one on the device; the 4 GB batch keeps the
    # first context, which already holds its workspace
    log("C4 concurrent: one context per database")
    want = []
    for r, nq, db, keys, qbuf, rbuf in groups:
        ctx.answer_batch_device(db, keys, qbuf, rbuf)
        ctx.synchronize()
        want.append(rbuf.download(np.zeros((nq, 2, N), dtype=np.uint64)))
    extra = [kspir_amd.Context(0) for _ in groups[:-1]]
    ctxs = extra + [ctx]
    for c, (r, nq, db, keys, qbuf, rbuf) in zip(ctxs, groups):   # warm-up
        c.answer_batch_device(db, keys, qbuf, rbuf)
    for c in ctxs:
        c.synchronize()
    for (r, nq, db, keys, qbuf, rbuf), w in zip(groups, want):
        if not np.array_equal(rbuf.download(np.zeros((nq, 2, N), dtype=np.uint64)), w):
            raise RuntimeError(f"C4 concurrent responses of r={r} differ from the sequential ones")
    t0 = time.perf_counter()
    for _ in range(steps):
        for c, (r, nq, db, keys, qbuf, rbuf) in zip(ctxs, groups):
            c.answer_batch_device(db, keys, qbuf, rbuf)
        for c in ctxs:
            c.synchronize()
    wall = (time.perf_counter() - t0) / steps
    out["mfma_concurrent"] = {"step_ms": round(wall * 1e3, 3), "qps": round(total_q / wall, 1), "contexts": len(ctxs),
                              "responses_equal_sequential": True}
    for c in extra:
        c.close()
    # one batch spanning the three databases: one launch per stage for all
    # 96 queries (kspir_answer_multi_device), checked against the per-database
    # batches, timed with HIP events
    log("C4 one launch per stage across the three databases")
    segs = [(db, keys) for r, nq, db, keys, qbuf, rbuf in groups]
    mq, mr = ctx.alloc(total_q * 4 * N * 8), ctx.alloc(total_q * 2 * N * 8)
    mq.upload(np.concatenate(host_q))
    ctx.answer_multi_device(segs, mq, mr)
    ctx.synchronize()
    if not np.array_equal(mr.download(np.zeros((total_q, 2, N), dtype=np.uint64)), np.concatenate(want)):
        raise RuntimeError("C4 responses of the batch spanning databases differ from the per-database batches")
    ms = ctx.time_answer_multi(segs, mq, mr, 1, steps) / steps
    out["mfma_one_launch"] = {"step_ms": round(ms, 3), "qps": round(total_q * 1000.0 / ms, 1),
                              "responses_equal_per_database": True}
    mq.free()
    mr.free()
    for r, nq, db, keys, qbuf, rbuf in groups:
        for k in keys:
            k.close()
        db.close()
        qbuf.free()
        rbuf.free()
    ctx.close()
    return out


def config_keygen(args, log):
    """client side on the GPU: secrets, queries and all keys of 64 clients
    at the 256 MB shape (r 16, N1 128), against the oracle's generation of
    one client on the host"""
    import kspir_amd
    r, n1, nc = 16, 128, 64
    ctx = kspir_amd.Context(0)
    seeds = list(range(1, nc + 1))
    tc = [(37 * i) % (N // 2) for i in range(nc)]
    tp = [i % r for i in range(nc)]
    ctx.client_generate(r, n1, seeds[:2], tc[:2], tp[:2])      # warm-up (tables, code)
    reps = []
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.client_generate(r, n1, seeds, tc, tp)
        reps.append(time.perf_counter() - t0)
    ctx.close()
    gpu_s = float(np.median(reps))
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle
        t0 = time.perf_counter()
        oracle.gen_client(r, n1, 1, tc[0], tp[0])
        cpu = time.perf_counter() - t0
    key_mb = (N + 4 * N + 8 * N + (n1 // 2) * 12 * N + (N // 2 // n1 - 1) * 4 * N + 4 * 8 * N) * 8 / 1e6
    return {"metric": "client key generation (secret, query, RGSW, BSGS and packing keys) per second",
            "value": round(nc / gpu_s, 2), "unit": "clients/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "uint64", "data": "synthetic seeds 1..64",
            "config": {"workload": f"{nc} clients at r={r}, N1={n1}, glibc rand() stream by jump-ahead on the GPU, "
                                   f"{key_mb:.1f} MB of key and query words per client returned to host memory",
                       "clients": nc},
            "ms_per_batch": round(gpu_s * 1e3, 2),
            "cpu_baseline": None if cpu is None else
            {"value": round(1.0 / cpu, 3), "unit": "clients/s", "cores": 1, "kind": "port",
             "sample": f"one client (seed 1) by the oracle on the host, {cpu * 1e3:.0f} ms"}}


def config_offline(args, log):
    """giant-step half of the offline key derivation (genAutoKeyFromOffline)
    for 64 clients at N1 = 64 (tests/test-two-steps-key.cpp's shape: a chain
    of 30 key switches by 5^N1, then 31 keys switched to crtMod), against the
    oracle's derivation for one client on the host"""
    import kspir_amd
    from oracle import oracle
    n1, nc = 64, 64
    log(f"offline keys for {nc} clients (oracle key generation on the host)")
    gens = [oracle.gen_offline_giant(n1, 100 + c) for c in range(nc)]
    off = np.stack([g[1] for g in gens])
    auto = np.stack([g[2] for g in gens])
    ctx = kspir_amd.Context(0)
    ctx.gen_giant_keys_from_offline(n1, off[:2], auto[:2], want_chain=False)   # warm-up
    reps = []
    for _ in range(3):
        t0 = time.perf_counter()
        _, giant = ctx.gen_giant_keys_from_offline(n1, off, auto, want_chain=False)
        reps.append(time.perf_counter() - t0)
    ctx.close()
    gpu_s = float(np.median(reps))
    cpu = None
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        _, want = oracle.giant_from_offline(n1, off[0], auto[0], want_chain=False)
        cpu = time.perf_counter() - t0
        assert np.array_equal(giant[0], want)
    n2 = N // 2 // n1
    return {"metric": "giant-step keys derived from offline keys per second",
            "value": round(nc / gpu_s, 2), "unit": "clients/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "uint32 RNS (q1, q2, auxMod)", "data": "synthetic keys from seeds 100..163",
            "config": {"workload": f"{nc} clients at N1={n1}: {n2 - 2} chained key switches by 5^N1 and {n2 - 1} "
                                   f"modulus switches per client; host keys in, {n2 - 1} giant keys per client "
                                   f"({(n2 - 1) * 6 * N * 8 / 1e6:.1f} MB) back to host memory, PCIe included",
                       "clients": nc},
            "ms_per_batch": round(gpu_s * 1e3, 2),
            "cpu_baseline": None if cpu is None else
            {"value": round(1.0 / cpu, 3), "unit": "clients/s", "cores": 1, "kind": "port",
             "sample": f"one client (seed 100) by the oracle on the host, {cpu * 1e3:.0f} ms, equal to the GPU's words"}}


def serve_rates(srv, targets, rates, dur, seed, log, tag, share=None):
    """Poisson arrivals at each offered rate for `dur` seconds through the
    server, each arrival sent to one of `targets` ((database, keys, queries)
    triples) drawn with probabilities `share` (default: all to the first);
    p50/p99 of the server's submit-to-response latency"""
    import threading
    res = {}
    rng = np.random.default_rng(seed)
    share = share or [1.0] + [0.0] * (len(targets) - 1)
    for rate in rates:
        log(f"C5 {tag}: {rate} queries/s offered for {dur} s")
        tickets, done_ms = [], []
        stop_flag = [False]

        def collect():
            i = 0
            while True:
                while i >= len(tickets):
                    if stop_flag[0] and i >= len(tickets):
                        return
                    time.sleep(0.0005)
                _, ms = srv.wait(tickets[i])
                done_ms.append(ms)
                i += 1

        th = threading.Thread(target=collect)
        th.start()
        b0 = srv.stats()
        t_start = time.perf_counter()
        t_next = t_start
        c = 0
        sent = [0] * len(targets)
        while t_next - t_start < dur:
            now = time.perf_counter()
            if now < t_next:
                time.sleep(t_next - now)
            g = int(rng.choice(len(targets), p=share))
            db, keys, queries = targets[g]
            tickets.append(srv.submit(db, keys[c % len(keys)], queries[c % len(queries)]))
            sent[g] += 1
            c += 1
            t_next += rng.exponential(1.0 / rate)
        stop_flag[0] = True
        th.join()
        wall = time.perf_counter() - t_start
        b1 = srv.stats()
        nb = b1["batches"] - b0["batches"]
        res[str(rate)] = {"offered_qps": rate, "submitted": c, "achieved_qps": round(c / wall, 1),
                          "latency_ms": {"p50": round(float(np.percentile(done_ms, 50)), 3),
                                         "p99": round(float(np.percentile(done_ms, 99)), 3)},
                          "batches": nb, "mean_batch": round(c / max(nb, 1), 2)}
        if len(targets) > 1:
            res[str(rate)]["submitted_per_database"] = sent
            res[str(rate)]["batches_spanning_databases"] = b1["spanning"] - b0["spanning"]
    return res


def config_c5(args, log):
    """C5: serving loop.  A resident 4 GB database (r=256) and, memory
    permitting, a 16 GB one (r=1024); 64 clients per database with their own
    keys; queries arrive as a Poisson process at each offered rate and go
    through the library's continuous-batching server (kspir_server_*): each
    database alone, with one lane and with two lanes (batches overlapping on
    the device), the 16 GB one also with graph capture and replay in the
    loop (option serve_graphs); then both databases resident together with
    the arrivals split evenly between them, one lane, with batches spanning
    the two databases (option serve_multi) and without; then the same over
    four 256 MB databases.  Latency is submit-to-response, measured by the
    server."""
    import kspir_amd
    n1 = 128
    ctx = kspir_amd.Context(0)
    perms = kspir_amd.compute_permutation_matrix(n1)
    rec = synthetic_records()
    out = {"config": "C5 serving: Poisson arrivals (5 s per offered rate per database alone, 3 s in the mixed runs), "
                     "continuous batching (batch cap 128 at 4 GB, 64 at 16 GB; admission "
                     "by each batch's device footprint; 1 lane, and 2 lanes whose batches overlap on the device), "
                     "64 clients with distinct keys per database, resident 4 GB (r=256) and 16 GB (r=1024) databases, "
                     "N1=128 (16 GB also with graph replay in the loop); then both resident with arrivals split "
                     "evenly, and four 256 MB databases with arrivals split evenly, batches spanning databases or not",
           "databases": {}}
    held = []
    for r, rates, dur, cap, lane_set in ((256, (200, 500, 1000, 2000), 5.0, 128, (1, 2)),
                                         (1024, (200, 500), 5.0, 64, (1,))):
        try:
            db = ctx.database_from_records(rec, r, n1, block_stride=0)
        except kspir_amd.KspirError as e:
            out["databases"][f"{r // 64} GB"] = {"skipped": str(e)}
            continue
        log(f"C5 database r={r}: 64 clients")
        keys, queries = [], []
        for c in range(64):
            query, baby, giant, packing, rgsw = synthetic_client(r, n1, 9000 + c)
            keys.append(ctx.keys(r, n1, baby, giant, packing, rgsw, perms))
            queries.append(query)
        by_lanes = {}
        # the 16 GB database also with graph capture and replay in the loop
        # (serve_graphs 1), against the default eager launches
        runs = [(lanes, 0) for lanes in lane_set] + ([(1, 1)] if r == 1024 else [])
        for lanes, graphs in runs:
            ctx.set("serve_graphs", graphs)
            srv = kspir_amd.Server(ctx, max_batch=cap, lanes=lanes)
            # warm-up: full batches on every lane (database limbs, workspaces at full size)
            for t in [srv.submit(db, keys[c % 64], queries[c % 64]) for c in range(cap * lanes)]:
                srv.wait(t)
            by_lanes[f"lanes_{lanes}" + ("_serve_graphs" if graphs else "")] = serve_rates(
                srv, [(db, keys, queries)], rates, dur, r, log, f"r={r} lanes={lanes} serve_graphs={graphs}")
            srv.close()
        ctx.set("serve_graphs", 0)
        out["databases"][f"{r // 64} GB"] = {"r": r, "batch_cap": cap,
                                            "per_query_footprint_MB": round(kspir_amd.query_footprint(r, n1) / 2**20, 1),
                                            **by_lanes}
        held.append((db, keys, queries))
    if len(held) == 2:
        mixed = {"share": [0.5, 0.5], "batch_cap": 64}
        for multi in (1, 0):
            ctx.set("serve_multi", multi)
            srv = kspir_amd.Server(ctx, max_batch=64, lanes=1)
            for t in [srv.submit(held[c % 2][0], held[c % 2][1][c % 64], held[c % 2][2][c % 64]) for c in range(64)]:
                srv.wait(t)
            mixed[f"serve_multi_{multi}"] = serve_rates(srv, held, (200, 500), 3.0, 77, log,
                                                        f"both databases serve_multi={multi}", share=[0.5, 0.5])
            srv.close()
        ctx.set("serve_multi", 0)
        out["both_resident"] = mixed
    for db, keys, _ in held:
        for k in keys:
            k.close()
        db.close()
    # four 256 MB databases (r=16) of distinct records, 32 clients each,
    # arrivals split evenly: batches spanning databases or not
    four = []
    for d in range(4):
        db = ctx.database_from_records(synthetic_records(100 + d), 16, n1, block_stride=0)
        keys, queries = [], []
        for c in range(32):
            query, baby, giant, packing, rgsw = synthetic_client(16, n1, 5000 + 100 * d + c)
            keys.append(ctx.keys(16, n1, baby, giant, packing, rgsw, perms))
            queries.append(query)
        four.append((db, keys, queries))
    eq = {"share": [0.25] * 4, "batch_cap": 64}
    for multi in (1, 0):
        ctx.set("serve_multi", multi)
        srv = kspir_amd.Server(ctx, max_batch=64, lanes=1)
        for t in [srv.submit(four[c % 4][0], four[c % 4][1][c % 32], four[c % 4][2][c % 32]) for c in range(128)]:
            srv.wait(t)
        eq[f"serve_multi_{multi}"] = serve_rates(srv, four, (4000, 12000), 3.0, 16, log,
                                                 f"four 256 MB databases serve_multi={multi}", share=[0.25] * 4)
        srv.close()
    ctx.set("serve_multi", 0)
    out["four_256MB_databases"] = eq
    for db, keys, _ in four:
        for k in keys:
            k.close()
        db.close()
    ctx.close()
    return out

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed batched answers")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=64, help="queries per step (distinct clients)")
    ap.add_argument("--db-mb", type=int, default=256, help="database size: 16 MB per block (r = MB/16)")
    ap.add_argument("--from-records", action="store_true",
                    help="build the database on the GPU from one synthetic block of records repeated (fast at "
                         "4 GB and up) instead of uploading uniform preprocessed words")
    ap.add_argument("--n1", type=int, default=128)
    ap.add_argument("--plain-steps", type=int, default=100)
    ap.add_argument("--lat-samples", type=int, default=1000, help="latency samples (plain and batched)")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch kernel by kernel instead of replaying hipGraphs")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default kernel shapes and paths")
    ap.add_argument("--config", default="c3", choices=("c3", "c2", "c4", "c5", "keygen", "offline"),
                    help="c3: the headline batched line (default); c2/c4/c5/keygen/offline: the other workloads, "
                         "one JSON line each")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the KsPIR answer here runs on a single MI355X "
              "(multi-GPU sharding by database block is a design note, not built).", file=sys.stderr)
        sys.exit(2)

    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)
    if args.config != "c3":
        fn = {"c2": config_c2, "c4": config_c4, "c5": lambda a, l: config_c5(a, l), "keygen": config_keygen,
              "offline": config_offline}[args.config]
        print(json.dumps(fn(args, log)), flush=True)
        return

    import kspir_amd
    r, n1, B = args.db_mb // 16, args.n1, args.batch
    n2 = N // 2 // n1
    log(f"synthetic database r={r} n1={n1}")
    ctx = kspir_amd.Context(0)
    if args.no_graphs:
        ctx.set("graphs", 0)
    if args.from_records:
        db = ctx.database_from_records(synthetic_records(), r, n1, block_stride=0)
    else:
        db = ctx.database(synthetic_db(r), r, n1)
    perms = kspir_amd.compute_permutation_matrix(n1)
    log(f"{B} synthetic clients")
    keys, queries = [], []
    for c in range(B):
        query, baby, giant, packing, rgsw = synthetic_client(r, n1, 1000 + c)
        keys.append(ctx.keys(r, n1, baby, giant, packing, rgsw, perms))
        queries.append(query)
    qbuf = ctx.alloc(B * 4 * N * 8)
    qbuf.upload(np.stack(queries))
    rbuf = ctx.alloc(B * 2 * N * 8)
    tuned = None
    if not args.no_autotune:
        log("autotuning kernel shapes and paths for this database")
        serve_b, rep = ctx.autotune(db, keys[:8], max_batch=B)
        tuned = {k: ctx.get(k) for k in ("db_waves", "mfma_min_batch", "mfma_max_chunk", "fused_min_batch",
                                          "fused_pack", "giant_wgs_per_cu", "giant_dma", "pack_fused_min_wgs")}
        tuned["serve_batch"] = serve_b
        tuned["report"] = rep

    # every response of one batched step equals its query's single answer
    log("checking one batched step against the single-query answers")
    ctx.answer_batch_device(db, keys, qbuf, rbuf)
    ctx.synchronize()
    batch_resp = rbuf.download(np.zeros((B, 2, N), dtype=np.uint64))
    q1buf, r1buf = ctx.alloc(4 * N * 8), ctx.alloc(2 * N * 8)
    for c in range(B):
        q1buf.upload(queries[c])
        ctx.answer_device(db, keys[c], q1buf, r1buf)
        ctx.synchronize()
        if not np.array_equal(r1buf.download(np.zeros((2, N), dtype=np.uint64)), batch_resp[c]):
            raise SystemExit(f"batched response {c} differs from its single-query answer")
    q1buf.free()
    r1buf.free()

    # plain single-query answer (client 0), one query in flight
    log("timing plain single-query answers")
    ctx.synchronize()
    plain_ms = ctx.time_answer(db, keys[0], qbuf, rbuf, 10, args.plain_steps) / args.plain_steps
    plain_lat = [ctx.time_answer(db, keys[0], qbuf, rbuf, 0, 1) for _ in range(args.lat_samples)]
    plain_eager_ms = None
    if not args.no_graphs:
        ctx.set("graphs", 0)
        plain_eager_ms = ctx.time_answer(db, keys[0], qbuf, rbuf, 10, args.plain_steps) / args.plain_steps
        ctx.set("graphs", 1)

    # batched answer: B queries of B clients per step
    log(f"timing batched answers (B={B})")
    total_ms = ctx.time_answer_batch(db, keys, qbuf, rbuf, args.warmup, args.steps)
    ms_per_step = total_ms / args.steps
    qps = B * 1000.0 / ms_per_step
    batch_lat = [ctx.time_answer_batch(db, keys, qbuf, rbuf, 0, 1) for _ in range(args.lat_samples)]
    # the batched database pass, kernel by kernel (HIP events): query limbs,
    # int8-MFMA products, move into the first-dimension layout
    db3 = ctx.time_db_pass_batch(db, keys, qbuf, 10) if B >= 2 else None

    # dominant single-query kernel: the database pass, timed alone with HIP
    # events on the context stream; algorithmic bytes = 8 B per word, r*N*N/2 words
    db_ms = ctx.time_db_pass(db, keys[0], qbuf, 20)
    db_bytes = 8 * r * N * (N // 2)
    achieved = db_bytes / (db_ms * 1e-3) / 1e9
    plain_qps = 1000.0 / plain_ms
    # what a plain 16-byte copy kernel reaches on this GPU (read + write bytes
    # over time); its HBM counters also calibrate the traffic figure
    log("timing a 256 MiB device copy")
    cbytes = 256 << 20
    csrc, cdst = ctx.alloc(cbytes), ctx.alloc(cbytes)
    copy_ms = ctx.time_copy(cdst, csrc, cbytes, 10)
    copy_gbps = 2 * cbytes / (copy_ms * 1e-3) / 1e9
    csrc.free()
    cdst.free()

    out = {
        "metric": "PIR queries/sec + DB GB/s at 256 MB, batched vs plain 1-query; p50/p99 latency",
        "value": round(qps, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(qps * README_SECONDS_256MB, 1) if args.db_mb == 256 and n1 == 128 else None,
        "dtype": "uint64",
        "data": "synthetic",
        "config": {
            "workload": f"batched online answer: {B} queries/step from {B} clients with distinct keys, "
                        f"one {args.db_mb} MB database ({r * N // 2} records x 8 KB), N=4096, r={r}, "
                        f"N1={n1}, N2={n2}",
            "records": r * N // 2, "record_bytes": 8192, "r": r, "N1": n1, "N2": n2, "queries_per_step": B,
            "database": "built on the GPU from one synthetic record block repeated" if args.from_records
                        else "uniform preprocessed words uploaded",
            "launch": "kernel by kernel" if args.no_graphs else "hipGraph replay",
            "checked": "every response of one batched step equals its query's single-query answer",
            "tuned": tuned,
        },
        "db_GBps_raw": round(r * N * (N // 2) * 2 * qps / 1e9, 2),
        "latency_ms": {"p50": round(float(np.percentile(batch_lat, 50)), 4),
                       "p99": round(float(np.percentile(batch_lat, 99)), 4), "samples": len(batch_lat)},
        "plain_single_query": {
            "qps": round(plain_qps, 2), "ms": round(plain_ms, 4),
            "db_GBps_raw": round(r * N * (N // 2) * 2 * plain_qps / 1e9, 2),
            "db_GBps_streamed": round(db_bytes * plain_qps / 1e9, 2),
            "ms_kernel_by_kernel": None if plain_eager_ms is None else round(plain_eager_ms, 4),
            "latency_ms": {"p50": round(float(np.percentile(plain_lat, 50)), 4),
                           "p99": round(float(np.percentile(plain_lat, 99)), 4), "samples": len(plain_lat)},
        },
        "batched_over_plain": round(qps / plain_qps, 2),
        "roofline_batched": batched_roofline(r, n1, B, db3, direct=bool(ctx.get("mfma_direct"))),
        "roofline": {"kernel": "k_db_product (single-query database pass)", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic_from_profile(),
                     "algorithmic_bytes": db_bytes, "avg_launch_ms": round(db_ms, 4),
                     "copy_kernel_GBps": round(copy_gbps, 1)},
    }
    for k in keys:
        k.close()
    db.close()
    qbuf.free()
    rbuf.free()
    ctx.close()
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"], out["cpu_baseline_all_cores"] = cpu_baseline(r, n1, args.cpu_reps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
