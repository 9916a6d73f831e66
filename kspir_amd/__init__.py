"""kspir_amd — MI355X-native KsPIR online server answer (ctypes over the C ABI).

The compute path is libkspir_amd.so (hand-written HIP kernels for gfx950)
built in-tree by kspir_amd/build.py.  There is no CPU fallback: importing
this package without the built library, or using it without a GPU, raises.
Host arrays are numpy uint64 / int32 in the reference's layouts (see
include/kspir_amd.h).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkspir_amd.so")
N = 4096

_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p


class KspirError(RuntimeError):
    """a library error; .code is the C-ABI return code (KSPIR_* in include/kspir_amd.h)"""
    code = None


# return codes (include/kspir_amd.h)
EINVAL, EHIP, ENOMEM, ECANCELED, EBUDGET, EBUSY, EINJECTED = -1, -2, -3, -4, -5, -6, -7


def _load():
    if not os.path.exists(LIB_PATH):
        raise KspirError(f"{LIB_PATH} is not built; run python kspir_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    L.kspir_last_error.restype = ctypes.c_char_p
    L.kspir_context_stream.restype = _vp
    L.kspir_context_stream.argtypes = [_vp]
    for name in ("kspir_context_create", "kspir_database_create", "kspir_keys_create", "kspir_answer",
                 "kspir_answer_device", "kspir_answer_inspect", "kspir_matrix_vector_mul_bsgs_rns_crt_large",
                 "kspir_external_product", "kspir_packing_rlwes", "kspir_device_alloc", "kspir_device_free",
                 "kspir_memcpy_h2d", "kspir_memcpy_d2h", "kspir_synchronize", "kspir_time_answer",
                 "kspir_time_db_pass", "kspir_test_ntt", "kspir_device_count",
                 "kspir_compute_permutation_matrix", "kspir_answer_batch", "kspir_answer_batch_device",
                 "kspir_time_answer_batch", "kspir_context_set", "kspir_time_copy",
                 "kspir_database_create_from_records", "kspir_database_tobsgsntt", "kspir_database_export",
                 "kspir_server_create", "kspir_server_submit", "kspir_server_wait", "kspir_server_stats",
                 "kspir_query_footprint_bytes", "kspir_decrypt_bsgs_total", "kspir_context_get",
                 "kspir_autotune", "kspir_client_generate", "kspir_client_draws", "kspir_client_encrypt",
                 "kspir_server_create_lanes", "kspir_gen_giant_keys_from_offline", "kspir_answer_device_noalloc",
                 "kspir_inverse_encode", "kspir_time_db_pass_batch", "kspir_answer_multi",
                 "kspir_answer_multi_device", "kspir_time_answer_multi", "kspir_database_compact",
                 "kspir_database_save", "kspir_database_load", "kspir_server_spanning_batches",
                 "kspir_server_cancel", "kspir_batch_footprint_bytes"):
        getattr(L, name).restype = ctypes.c_int
    L.kspir_autotune.argtypes = [_vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                 ctypes.POINTER(ctypes.c_int32), ctypes.c_char_p, ctypes.c_size_t]
    L.kspir_context_destroy.argtypes = [_vp]
    L.kspir_database_destroy.argtypes = [_vp]
    L.kspir_keys_destroy.argtypes = [_vp]
    L.kspir_server_destroy.argtypes = [_vp]
    return L


lib = _load()

# every symbol include/kspir_amd.h and include/kspir_amd_runtime.h declare
EXPORTS = (
    "kspir_last_error", "kspir_context_create", "kspir_context_destroy", "kspir_database_create",
    "kspir_database_destroy", "kspir_keys_create", "kspir_keys_destroy", "kspir_answer",
    "kspir_answer_device", "kspir_matrix_vector_mul_bsgs_rns_crt_large", "kspir_external_product",
    "kspir_packing_rlwes", "kspir_answer_inspect", "kspir_compute_permutation_matrix",
    "kspir_answer_batch", "kspir_answer_batch_device", "kspir_time_answer_batch", "kspir_context_set",
    "kspir_device_count", "kspir_device_alloc", "kspir_device_free", "kspir_memcpy_h2d", "kspir_memcpy_d2h",
    "kspir_synchronize", "kspir_context_stream", "kspir_time_answer", "kspir_time_db_pass", "kspir_test_ntt",
    "kspir_time_copy", "kspir_database_create_from_records", "kspir_database_tobsgsntt", "kspir_database_export",
    "kspir_server_create", "kspir_server_submit", "kspir_server_wait", "kspir_server_stats", "kspir_server_destroy",
    "kspir_query_footprint_bytes", "kspir_decrypt_bsgs_total", "kspir_context_get", "kspir_autotune",
    "kspir_client_generate", "kspir_client_draws", "kspir_client_encrypt", "kspir_server_create_lanes",
    "kspir_gen_giant_keys_from_offline", "kspir_answer_device_noalloc", "kspir_inverse_encode",
    "kspir_time_db_pass_batch", "kspir_answer_multi", "kspir_answer_multi_device", "kspir_time_answer_multi",
    "kspir_database_compact", "kspir_database_save", "kspir_database_load", "kspir_server_spanning_batches",
    "kspir_server_cancel", "kspir_batch_footprint_bytes",
)


def _check(rc):
    if rc != 0:
        e = KspirError(f"kspir error {rc}: {lib.kspir_last_error().decode()}")
        e.code = rc
        raise e


def _p64(a):
    if a is None:
        return ctypes.cast(None, _u64p)
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], "expected C-contiguous uint64"
    return a.ctypes.data_as(_u64p)


def _p32(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], "expected C-contiguous int32"
    return a.ctypes.data_as(_i32p)


def compute_permutation_matrix(n1):
    """utils.cpp:1591 compute_permutation_matrix(n1, N), host-side."""
    out = np.zeros((n1, N), dtype=np.int32)
    _check(lib.kspir_compute_permutation_matrix(_p32(out), n1))
    return out


def device_count():
    n = ctypes.c_int(0)
    _check(lib.kspir_device_count(ctypes.byref(n)))
    return n.value


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx, self.nbytes = ctx, nbytes
        p = _vp()
        _check(lib.kspir_device_alloc(ctx.h, ctypes.c_size_t(nbytes), ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib.kspir_memcpy_h2d(self.ctx.h, _vp(self.ptr), arr.ctypes.data_as(_vp), ctypes.c_size_t(arr.nbytes)))

    def download(self, arr):
        assert arr.flags["C_CONTIGUOUS"] and arr.nbytes <= self.nbytes
        _check(lib.kspir_memcpy_d2h(self.ctx.h, arr.ctypes.data_as(_vp), _vp(self.ptr), ctypes.c_size_t(arr.nbytes)))
        return arr

    def free(self):
        if self.ptr:
            _check(lib.kspir_device_free(self.ctx.h, _vp(self.ptr)))
            self.ptr = None


class Database:
    """Preprocessed database resident in HBM (database_tocrt words, r blocks)."""

    def __init__(self, ctx, datacrt, r, n1, on_device=False, records=False, block_stride=None):
        self.ctx, self.r, self.n1 = ctx, r, n1
        h = _vp()
        ptr = ctypes.cast(_vp(datacrt), _u64p) if on_device else _p64(datacrt)
        if records:
            stride = N * N // 2 if block_stride is None else block_stride
            _check(lib.kspir_database_create_from_records(ctx.h, ptr, ctypes.c_int64(stride), r, n1, int(on_device),
                                                          ctypes.byref(h)))
        else:
            _check(lib.kspir_database_create(ctx.h, ptr, r, n1, int(on_device), ctypes.byref(h)))
        self.h = h

    def export(self):
        """the words in the reference's datacrt layout, [r][N*N/2]"""
        out = np.zeros((self.r, N * N // 2), dtype=np.uint64)
        _check(lib.kspir_database_export(self.h, _p64(out)))
        return out

    def save(self, path):
        """write the persisted form (header + datacrt words) to `path`"""
        _check(lib.kspir_database_save(self.h, os.fsencode(path)))

    def compact(self):
        """keep only the int8 limbs resident (frees the u64 words; every answer
        then takes the MFMA pass)"""
        _check(lib.kspir_database_compact(self.h))

    def close(self):
        if self.h:
            lib.kspir_database_destroy(self.h)
            self.h = None


def _database_load(ctx, path):
    import struct
    with open(path, "rb") as f:
        head = f.read(24)
    h = _vp()
    _check(lib.kspir_database_load(ctx.h, os.fsencode(path), ctypes.byref(h)))
    db = Database.__new__(Database)
    db.ctx, db.h = ctx, h
    db.r, db.n1 = struct.unpack_from("<II", head, 16)
    return db


class Keys:
    """Per-client evaluation keys resident in HBM."""

    def __init__(self, ctx, r, n1, baby, giant, packing, rgsw, perms):
        self.ctx, self.r, self.n1 = ctx, r, n1
        h = _vp()
        giant = giant if giant is not None and giant.size else None
        packing = packing if packing is not None and packing.size else None
        _check(lib.kspir_keys_create(ctx.h, r, n1, _p64(baby), _p64(giant), _p64(packing), _p64(rgsw),
                                     _p32(perms), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib.kspir_keys_destroy(self.h)
            self.h = None


def query_footprint(r, n1):
    """device bytes one query of shape (r, N1) adds to a batch"""
    out = ctypes.c_size_t(0)
    _check(lib.kspir_query_footprint_bytes(r, n1, ctypes.byref(out)))
    return out.value


def batch_footprint(r, n1, nq, ctx=None):
    """device bytes of a batch of nq queries of shape (r, N1) under ctx's
    options (None: the defaults), as the serving loop admits it"""
    out = ctypes.c_size_t(0)
    _check(lib.kspir_batch_footprint_bytes(ctx.h if ctx is not None else None, r, n1, nq, ctypes.byref(out)))
    return out.value


class Server:
    """Continuous-batching serving loop on a context (see kspir_server_* in
    include/kspir_amd.h).  While it runs, the context is driven by its worker
    thread.  submit() returns a ticket; wait(ticket) returns (response,
    latency_ms)."""

    def __init__(self, ctx, max_batch=64, budget_bytes=0, lanes=1):
        self.ctx = ctx
        h = _vp()
        _check(lib.kspir_server_create_lanes(ctx.h, max_batch, ctypes.c_size_t(budget_bytes), int(lanes),
                                             ctypes.byref(h)))
        self.h = h
        self._live = {}

    def submit(self, db, keys, query):
        q = np.ascontiguousarray(query, dtype=np.uint64)
        resp = np.zeros((2, N), dtype=np.uint64)
        t = ctypes.c_uint64(0)
        _check(lib.kspir_server_submit(self.h, db.h, keys.h, _p64(q), _p64(resp), ctypes.byref(t)))
        self._live[t.value] = (q, resp)      # kept alive until wait()
        return t.value

    def wait(self, ticket):
        ms = ctypes.c_double(0)
        rc = lib.kspir_server_wait(self.h, ctypes.c_uint64(ticket), ctypes.byref(ms))
        # the library forgets the ticket whatever its result (an unknown
        # ticket was never kept here)
        q, resp = self._live.pop(ticket, (None, None))
        _check(rc)
        return resp, ms.value

    def cancel(self, ticket):
        """withdraw a request no batch has taken yet (its wait then raises
        KspirError with code ECANCELED); raises with EBUSY when too late"""
        _check(lib.kspir_server_cancel(self.h, ctypes.c_uint64(ticket)))

    def stats(self):
        b, a, p = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(lib.kspir_server_stats(self.h, ctypes.byref(b), ctypes.byref(a), ctypes.byref(p)))
        sp = ctypes.c_uint64(0)
        _check(lib.kspir_server_spanning_batches(self.h, ctypes.byref(sp)))
        return {"batches": b.value, "answered": a.value, "pending": p.value, "spanning": sp.value}

    def close(self):
        if self.h:
            lib.kspir_server_destroy(self.h)
            self.h = None


class Context:
    def __init__(self, device=0):
        h = _vp()
        _check(lib.kspir_context_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib.kspir_context_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set(self, key, value):
        _check(lib.kspir_context_set(self.h, key.encode(), ctypes.c_int64(value)))

    def client_generate(self, r, n1, seeds, target_cols, target_packings):
        """secrets, queries and keys of len(seeds) clients on the GPU, each as
        after srand(seed) in test-pir; returns a dict of arrays in the
        reference layouts (leading axis: client)"""
        nc = len(seeds)
        n2, lr = N // 2 // n1, int(r).bit_length() - 1
        out = {"secret": np.zeros((nc, N), np.uint64), "query": np.zeros((nc, 2, 2, N), np.uint64),
               "rgsw": np.zeros((nc, 4, 2, N), np.uint64), "baby": np.zeros((nc, n1 // 2, 6, 2, N), np.uint64),
               "giant": np.zeros((nc, max(n2 - 1, 0), 2, 2, N), np.uint64),
               "packing": np.zeros((nc, lr, 4, 2, N), np.uint64)}
        sd = np.ascontiguousarray(seeds, dtype=np.uint32)
        tc = np.ascontiguousarray(target_cols, dtype=np.int32)
        tp = np.ascontiguousarray(target_packings, dtype=np.int32)
        _check(lib.kspir_client_generate(self.h, int(r), int(n1), nc, sd.ctypes.data_as(_vp), tc.ctypes.data_as(_vp),
                                         tp.ctypes.data_as(_vp), *[_p64(out[k]) for k in
                                                                   ("secret", "query", "rgsw", "baby", "giant",
                                                                    "packing")]))
        return out

    def get(self, key):
        v = ctypes.c_int64(0)
        _check(lib.kspir_context_get(self.h, key.encode(), ctypes.byref(v)))
        return v.value

    def autotune(self, db, keys_list, max_batch=64, latency_budget_ms=0.0):
        """time the pipeline for this database and set the tuning options;
        returns (serving batch size, report dict of every timing)"""
        arr = (_vp * len(keys_list))(*[k.h.value for k in keys_list])
        sb = ctypes.c_int32(0)
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib.kspir_autotune(self.h, db.h, ctypes.cast(arr, _vp), len(keys_list), int(max_batch),
                                  float(latency_budget_ms), ctypes.byref(sb), buf, ctypes.c_size_t(len(buf))))
        import json
        return sb.value, json.loads(buf.value.decode())

    @property
    def stream(self):
        return lib.kspir_context_stream(self.h)

    def database(self, datacrt, r, n1, on_device=False):
        return Database(self, datacrt, r, n1, on_device)

    def database_load(self, path):
        """a database written by Database.save (header checked, words uploaded)"""
        return _database_load(self, path)

    def database_from_records(self, records, r, n1, block_stride=None, on_device=False):
        """preprocess records ([r][N][N/2] values < p, the sample_database_bsgs
        layout) on the GPU into a resident database; block_stride 0 repeats
        the first block"""
        if not on_device:
            records = np.ascontiguousarray(records, dtype=np.uint64)
        return Database(self, records, r, n1, on_device, records=True, block_stride=block_stride)

    def inverse_encode(self, columns):
        """inverse_encode of record columns on the GPU: [count][N] values < p"""
        out = np.ascontiguousarray(columns, dtype=np.uint64).reshape(-1, N).copy()
        _check(lib.kspir_inverse_encode(self.h, _p64(out), out.shape[0]))
        return out.reshape(np.shape(columns))

    def database_tobsgsntt(self, data, n1):
        """database_tobsgsntt of one block on the GPU: [N][N/2] -> [N/2][N] mod Q"""
        out = np.zeros((N // 2, N), dtype=np.uint64)
        _check(lib.kspir_database_tobsgsntt(self.h, _p64(np.ascontiguousarray(data, dtype=np.uint64)), n1,
                                            _p64(out)))
        return out

    def keys(self, r, n1, baby, giant, packing, rgsw, perms):
        return Keys(self, r, n1, baby, giant, packing, rgsw, perms)

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def synchronize(self):
        _check(lib.kspir_synchronize(self.h))

    def answer(self, db, keys, query):
        resp = np.zeros((2, N), dtype=np.uint64)
        _check(lib.kspir_answer(self.h, db.h, keys.h, _p64(np.ascontiguousarray(query)), _p64(resp)))
        return resp

    def answer_device(self, db, keys, query_buf, resp_buf, stream=None):
        _check(lib.kspir_answer_device(self.h, db.h, keys.h, ctypes.cast(_vp(query_buf.ptr), _u64p),
                                       ctypes.cast(_vp(resp_buf.ptr), _u64p), _vp(stream)))

    def answer_device_noalloc(self, db, keys, query_buf, resp_buf, stream=None):
        """kspir_answer_device_noalloc: the same answer on a fixed per-GPU
        scratch array, with no device allocation (asynchronous)"""
        _check(lib.kspir_answer_device_noalloc(self.h, db.h, keys.h, ctypes.cast(_vp(query_buf.ptr), _u64p),
                                               ctypes.cast(_vp(resp_buf.ptr), _u64p), _vp(stream)))

    def answer_batch(self, db, keys_list, queries):
        """queries [nq][2][2][N] with one Keys per query -> responses [nq][2][N]"""
        nq = len(keys_list)
        q = np.ascontiguousarray(queries, dtype=np.uint64)
        assert q.shape == (nq, 2, 2, N)
        arr = (_vp * nq)(*[k.h.value for k in keys_list])
        resp = np.zeros((nq, 2, N), dtype=np.uint64)
        _check(lib.kspir_answer_batch(self.h, db.h, arr, nq, _p64(q), _p64(resp)))
        return resp

    def answer_batch_device(self, db, keys_list, queries_buf, resp_buf, stream=None):
        """enqueue a batch on device buffers (asynchronous; synchronize() to wait)"""
        nq = len(keys_list)
        arr = (_vp * nq)(*[k.h.value for k in keys_list])
        _check(lib.kspir_answer_batch_device(self.h, db.h, arr, nq, ctypes.cast(_vp(queries_buf.ptr), _u64p),
                                             ctypes.cast(_vp(resp_buf.ptr), _u64p), _vp(stream)))

    def time_answer_batch(self, db, keys_list, queries_buf, resp_buf, warmup, iters):
        nq = len(keys_list)
        arr = (_vp * nq)(*[k.h.value for k in keys_list])
        ms = ctypes.c_float(0)
        _check(lib.kspir_time_answer_batch(self.h, db.h, arr, nq, ctypes.cast(_vp(queries_buf.ptr), _u64p),
                                           ctypes.cast(_vp(resp_buf.ptr), _u64p), warmup, iters, ctypes.byref(ms)))
        return ms.value

    @staticmethod
    def _multi_args(segments):
        """segments: [(Database, [Keys, ...]), ...] -> ctypes arrays"""
        dbs = (_vp * len(segments))(*[db.h.value for db, _ in segments])
        counts = (ctypes.c_int32 * len(segments))(*[len(ks) for _, ks in segments])
        keys = [k for _, ks in segments for k in ks]
        arr = (_vp * len(keys))(*[k.h.value for k in keys])
        return len(segments), dbs, counts, arr, len(keys)

    def answer_multi(self, segments, queries):
        """one batch spanning databases: segments [(Database, [Keys per query])],
        queries [sum nq][2][2][N] in segment order -> responses [sum nq][2][N]"""
        nseg, dbs, counts, arr, nq = self._multi_args(segments)
        q = np.ascontiguousarray(queries, dtype=np.uint64)
        assert q.shape == (nq, 2, 2, N)
        resp = np.zeros((nq, 2, N), dtype=np.uint64)
        _check(lib.kspir_answer_multi(self.h, nseg, dbs, counts, arr, _p64(q), _p64(resp)))
        return resp

    def answer_multi_device(self, segments, queries_buf, resp_buf, stream=None):
        nseg, dbs, counts, arr, nq = self._multi_args(segments)
        _check(lib.kspir_answer_multi_device(self.h, nseg, dbs, counts, arr, ctypes.cast(_vp(queries_buf.ptr), _u64p),
                                             ctypes.cast(_vp(resp_buf.ptr), _u64p), _vp(stream)))

    def time_answer_multi(self, segments, queries_buf, resp_buf, warmup, iters):
        nseg, dbs, counts, arr, nq = self._multi_args(segments)
        ms = ctypes.c_float(0)
        _check(lib.kspir_time_answer_multi(self.h, nseg, dbs, counts, arr, ctypes.cast(_vp(queries_buf.ptr), _u64p),
                                           ctypes.cast(_vp(resp_buf.ptr), _u64p), warmup, iters, ctypes.byref(ms)))
        return ms.value

    def answer_inspect(self, db, keys, query):
        r, n1 = db.r, db.n1
        resp = np.zeros((2, N), dtype=np.uint64)
        cbuf = np.zeros((N, n1, 2), dtype=np.uint64)
        fd = np.zeros((r, 2, N), dtype=np.uint64)
        ext = np.zeros((r, 2, N), dtype=np.uint64)
        _check(lib.kspir_answer_inspect(self.h, db.h, keys.h, _p64(np.ascontiguousarray(query)), _p64(resp),
                                        _p64(cbuf), _p64(fd), _p64(ext)))
        return resp, {"cipherbuf": cbuf, "firstdim": fd, "ext": ext}

    def matrix_vector_mul(self, db, keys, query):
        """matrix_vector_mul_bsgs_rns_crt_large: returns (result [r][2][N], mutated query)."""
        q = np.ascontiguousarray(query, dtype=np.uint64).copy()
        res = np.zeros((db.r, 2, N), dtype=np.uint64)
        _check(lib.kspir_matrix_vector_mul_bsgs_rns_crt_large(self.h, db.h, keys.h, _p64(q), _p64(res)))
        return res, q

    def external_product(self, keys, result, first_dim, first_dim_is_ntt):
        res = np.ascontiguousarray(result, dtype=np.uint64).copy()
        fd = np.ascontiguousarray(first_dim, dtype=np.uint64).copy()
        _check(lib.kspir_external_product(self.h, keys.h, _p64(res), _p64(fd), int(first_dim_is_ntt)))
        return res, fd

    def packing_rlwes(self, keys, rlwes, rlwes_are_ntt=False):
        rl = np.ascontiguousarray(rlwes, dtype=np.uint64).copy()
        res = np.zeros((2, N), dtype=np.uint64)
        _check(lib.kspir_packing_rlwes(self.h, keys.h, _p64(res), _p64(rl), int(rlwes_are_ntt)))
        return res

    def time_answer(self, db, keys, query_buf, resp_buf, warmup, iters):
        ms = ctypes.c_float(0)
        _check(lib.kspir_time_answer(self.h, db.h, keys.h, ctypes.cast(_vp(query_buf.ptr), _u64p),
                                     ctypes.cast(_vp(resp_buf.ptr), _u64p), warmup, iters, ctypes.byref(ms)))
        return ms.value

    def time_db_pass(self, db, keys, query_buf, iters):
        ms = ctypes.c_float(0)
        _check(lib.kspir_time_db_pass(self.h, db.h, keys.h, ctypes.cast(_vp(query_buf.ptr), _u64p), iters,
                                      ctypes.byref(ms)))
        return ms.value

    def time_db_pass_batch(self, db, keys_list, queries_buf, iters):
        """average ms per launch of the batched database pass's three kernels:
        (query limbs, MFMA products, move into the first-dimension layout)"""
        nq = len(keys_list)
        arr = (_vp * nq)(*[k.h.value for k in keys_list])
        ms = (ctypes.c_float * 3)()
        _check(lib.kspir_time_db_pass_batch(self.h, db.h, arr, nq, ctypes.cast(_vp(queries_buf.ptr), _u64p), iters,
                                            ms))
        return list(ms)

    def time_copy(self, dst_buf, src_buf, nbytes, iters):
        """average ms of a plain 16-byte-per-lane copy of nbytes (device buffers)"""
        ms = ctypes.c_float(0)
        _check(lib.kspir_time_copy(self.h, _vp(dst_buf.ptr), _vp(src_buf.ptr), ctypes.c_size_t(nbytes), iters,
                                   ctypes.byref(ms)))
        return ms.value

    def gen_giant_keys_from_offline(self, n1, offline2, auto2, want_chain=True):
        """giant-step half of genAutoKeyFromOffline for nc clients on the GPU.
        offline2 [nc][10][2][N], auto2 [nc][6][2][N] (NTT form) -> (chain
        [nc][N2-2][6][2][N] or None, giant [nc][N2-1][3][2][N])"""
        off = np.ascontiguousarray(offline2, dtype=np.uint64).reshape(-1, 10, 2, N)
        au = np.ascontiguousarray(auto2, dtype=np.uint64).reshape(-1, 6, 2, N)
        nc = off.shape[0]
        assert au.shape[0] == nc
        n2 = N // 2 // n1
        giant = np.zeros((nc, max(n2 - 1, 0), 3, 2, N), dtype=np.uint64)
        chain = np.zeros((nc, max(n2 - 2, 0), 6, 2, N), dtype=np.uint64) if want_chain else None
        _check(lib.kspir_gen_giant_keys_from_offline(self.h, int(n1), int(nc), _p64(off), _p64(au),
                                                     _p64(chain) if chain is not None and chain.size else None,
                                                     _p64(giant)))
        return chain, giant

    def decrypt_bsgs_total(self, r, secrets, responses):
        """recover on the GPU: secrets [nq][N], responses [nq][2][N] -> messages [nq][N]"""
        sk = np.ascontiguousarray(secrets, dtype=np.uint64).reshape(-1, N)
        rs = np.ascontiguousarray(responses, dtype=np.uint64).reshape(-1, 2, N)
        assert sk.shape[0] == rs.shape[0]
        out = np.zeros((sk.shape[0], N), dtype=np.uint64)
        _check(lib.kspir_decrypt_bsgs_total(self.h, r, sk.shape[0], _p64(sk), _p64(rs), _p64(out)))
        return out

    def test_ntt(self, polys, prime, inverse):
        x = np.ascontiguousarray(polys, dtype=np.uint32).copy()
        _check(lib.kspir_test_ntt(self.h, x.ctypes.data_as(_u32p), x.shape[0], prime, int(inverse)))
        return x
