/*
 * kspir_amd.h — C ABI of the MI355X (gfx950) KsPIR online-answer library.
 *
 * The library replaces the online server response of the reference
 * (tests/test-pir.cpp:107-116): matrix_vector_mul_bsgs_rns_crt_large
 * (src/twosteps.h:55), externalProduct (src/lwe.h:292) and packingRLWEs
 * (src/answer.h:79).  Arguments are the reference's own data, flattened:
 *
 *   N = 4096, Q = 268369921 * 249561089 (crtMod), bsMod = 16760833.
 *   RLWE           uint64_t [2][N]            a then b, words mod Q
 *   query          uint64_t [2][2][N]         input[0] (mod Q) a,b then input[1] (mod bsMod) a,b,
 *                                             coefficient form (query_bsgs_rns, twosteps.cpp:279)
 *   database       uint64_t [r][N][N2][N1]    database_tocrt words (crt.cpp:859-885):
 *                                             (x mod q1) | (x mod q2) << 32
 *   baby keys      uint64_t [N1/2][6][2][N]   AutoKeyBSGSRNS::keyMap.at(5^i), i = 1..N1/2
 *                                             (rows 0..2 mod Q, rows 3..5 mod bsMod), NTT form
 *   giant keys     uint64_t [N2-1][2][2][N]   AutoKeyBSGSRNS::keyMap.at(5^(N1*i)), i = 1..N2-1
 *   packing keys   uint64_t [log2 r][4][2][N] AutoKey::keyMap.at(N/i + 1), i = 1,2,4,..,r/2
 *   rgsw           uint64_t [4][2][N]         RGSWCiphertext rows (a[k], b[k]), NTT form
 *   permutations   int32_t  [N1][N]           compute_permutation_matrix(N1, N) (utils.cpp:1591)
 *
 * Every value is exact modular integer arithmetic; results are word-for-word
 * those of the reference algorithm on the same inputs.  Functions return 0 on
 * success or a negative KSPIR_E* code; kspir_last_error() describes the last
 * failure on the calling thread.  A context is bound to one device and one
 * HIP stream and is not safe for concurrent use from several threads.
 * Database and key handles may be used by any context of their device.
 */
#ifndef KSPIR_AMD_H
#define KSPIR_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KSPIR_OK 0
#define KSPIR_EINVAL (-1)   /* bad argument or shape */
#define KSPIR_EHIP (-2)     /* HIP runtime error (message carries hipGetErrorString) */
#define KSPIR_ENOMEM (-3)   /* device allocation failed */
/* serving loop (kspir_server_*) */
#define KSPIR_ECANCELED (-4) /* the request was cancelled before it ran */
#define KSPIR_EBUDGET (-5)   /* one query of this shape needs more device workspace than a lane's budget */
#define KSPIR_EBUSY (-6)     /* the request is already running or answered: too late to cancel */
#define KSPIR_EINJECTED (-7) /* failure injected by the context option serve_fail_batches (failure-path tests) */

#define KSPIR_N 4096

typedef struct kspir_context kspir_context;
typedef struct kspir_database kspir_database;
typedef struct kspir_keys kspir_keys;

const char* kspir_last_error(void);

/* compute_permutation_matrix (utils.h:129, utils.cpp:1591-1619): out [n1][N],
 * row i maps NTT slots so that NTT(tau_{5^i}(a))[s] = NTT(a)[out[i][s]];
 * n1 in [1, N/2].  Host-only closed form; needs no device. */
int kspir_compute_permutation_matrix(int32_t* out, int32_t n1);

/* device context: twiddle tables, stream and scratch for one device */
int kspir_context_create(int device, kspir_context** out);
void kspir_context_destroy(kspir_context* ctx);

/* Preprocessed database, resident in HBM for the lifetime of the handle.
 * Replaces the `data` argument of matrix_vector_mul_bsgs_rns_crt_large
 * (twosteps.h:55, twosteps.cpp:1527); r blocks of N*N/2 words each.
 * `datacrt` is a host pointer, or a device pointer when on_device != 0.
 * r: a power of two <= N; N1: a power of two in [32, N/2] (N2 = N/2/N1). */
int kspir_database_create(kspir_context* ctx, const uint64_t* datacrt, int32_t r, int32_t n1,
                          int on_device, kspir_database** out);
void kspir_database_destroy(kspir_database* db);

/* No counterpart in the reference (a resident-memory choice).  A database
 * that has served a batch holds two forms in HBM: its u64 words (4 B per
 * residue, read by the single-query word pass) and its 7-bit int8 limbs
 * (4 B per residue, read by the int8-MFMA pass).  kspir_database_compact
 * builds the limbs if needed and frees the words, halving the footprint.
 * Afterwards every answer takes the MFMA pass (one query: 0.62 ms against
 * 0.50 ms at 256 MB, responses unchanged), kspir_database_export rebuilds
 * the words from the limbs, and kspir_answer_device_noalloc and
 * kspir_time_db_pass (hence kspir_autotune: tune first) reject the
 * database.  Only shapes with an MFMA pass (N1 64..256, r*N2 % 16 == 0) can
 * be compacted; no answer on this database may be in flight.  Idempotent. */
int kspir_database_compact(kspir_database* db);

/* Server preprocessing on the GPU: database_tobsgsntt (twosteps.h:22,
 * twosteps.cpp:42-179) followed by database_tocrt (crt.h:133,
 * crt.cpp:859-885), straight into a resident database handle.  records: r
 * blocks in the layout sample_database_bsgs fills (twosteps.cpp:23-39), N
 * rows of N/2 values < p per block; block b starts at records + b *
 * block_stride (block_stride = N*N/2, or 0 to repeat one block).  Host
 * pointer, or device pointer when on_device != 0.  The words equal those of
 * database_tobsgsntt + database_tocrt on the CPU. */
int kspir_database_create_from_records(kspir_context* ctx, const uint64_t* records, int64_t block_stride,
                                       int32_t r, int32_t n1, int on_device, kspir_database** out);

/* database_tobsgsntt (twosteps.cpp:42-179) for one block on the GPU: data
 * [N][N/2] record values < p (host), result [N/2][N] values mod Q in NTT
 * form (host), as the reference returns them. */
int kspir_database_tobsgsntt(kspir_context* ctx, const uint64_t* data, int32_t n1, uint64_t* result);

/* inverse_encode (twosteps.h:22, twosteps.cpp:182-198) of `count` record
 * columns on the GPU: messages [count][N] values < p = 65537 (host), each
 * replaced by its forward NTT mod p (HEXL's NTT(N, p), minimal root) in the
 * slot order of compute_query_decode (utils.cpp:1677-1685).  Values >= p
 * are rejected with KSPIR_EINVAL. */
int kspir_inverse_encode(kspir_context* ctx, uint64_t* messages, int32_t count);

/* Copy a resident database back to the host in the reference's datacrt
 * layout (crt.cpp:859-885: per block [N][N2][N1] words): the persisted form
 * that kspir_database_create loads again. */
int kspir_database_export(const kspir_database* db, uint64_t* datacrt);

/* No counterpart in the reference (which keeps the preprocessed database in
 * memory only).  The persisted form as a file: a 64-byte header (magic
 * "KSPIRDB", format version 1, N, r, N1, words per block, FNV-1a 64 checksum
 * of the words) followed by the kspir_database_export words.
 * kspir_database_load rejects a file whose magic, version, N or block size
 * differ, whose length does not match its r, or whose checksum fails, and
 * then uploads it as kspir_database_create does (r and N1 from the header). */
int kspir_database_save(const kspir_database* db, const char* path);
int kspir_database_load(kspir_context* ctx, const char* path, kspir_database** out);

/* Per-client evaluation keys, resident in HBM.  Replaces the autoKey and
 * permutations arguments of matrix_vector_mul_bsgs_rns_crt_large
 * (twosteps.cpp:1527), the query argument of externalProduct (lwe.cpp:761)
 * and the autokey argument of packingRLWEs (answer.cpp:800).  Host pointers.
 * Every part is optional (NULL) and checked by the call that needs it:
 * baby + permutations (+ giant when N2 > 1) for the first dimension, rgsw
 * for the external product, packing (r > 1) for packing. */
int kspir_keys_create(kspir_context* ctx, int32_t r, int32_t n1, const uint64_t* baby,
                      const uint64_t* giant, const uint64_t* packing, const uint64_t* rgsw,
                      const int32_t* permutations, kspir_keys** out);
void kspir_keys_destroy(kspir_keys* keys);

/* Online answer of test-pir.cpp:107-116: first dimension over the r database
 * blocks, r external products, packing.  query (host) -> response (host),
 * coefficient form.  The query is not modified. */
int kspir_answer(kspir_context* ctx, const kspir_database* db, const kspir_keys* keys,
                 const uint64_t* query, uint64_t* response);

/* Same, with query and response in device memory, enqueued on `stream`
 * (a hipStream_t; NULL selects the context's stream).  Asynchronous. */
int kspir_answer_device(kspir_context* ctx, const kspir_database* db, const kspir_keys* keys,
                        const uint64_t* query_dev, uint64_t* response_dev, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): kspir_answer_device with the same arguments and the
 * same response words, but it calls no hipMalloc, hipFree or allocating
 * library routine.  Its workspace is a fixed 64 MiB device array per GPU
 * (held from context creation until the process ends), so shapes whose
 * single-query workspace exceeds it are rejected with KSPIR_EINVAL: r 16 at
 * N1 128 (the 256 MB test-pir shape) and r 16 at N1 256 fit, r 32 at N1 128
 * does not.  Launches are eager (no graph) and use the VALU database pass.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (two streams, two contexts or two threads without ordering between
 * them) race and give undefined responses; order them on one stream or
 * synchronize between them.  Asynchronous, like kspir_answer_device. */
int kspir_answer_device_noalloc(kspir_context* ctx, const kspir_database* db, const kspir_keys* keys,
                                const uint64_t* query_dev, uint64_t* response_dev, void* stream);

/* Batched online answer: nq queries, each with its own client keys (all for
 * this database's r and N1), against one resident database.  Every database
 * tile read from HBM serves the whole batch.  queries [nq][2][2][N] ->
 * responses [nq][2][N]; response i equals kspir_answer on query i alone.
 * No counterpart in the reference, which answers one query per call. */
int kspir_answer_batch(kspir_context* ctx, const kspir_database* db, const kspir_keys* const* keys,
                       int32_t nq, const uint64_t* queries, uint64_t* responses);
int kspir_answer_batch_device(kspir_context* ctx, const kspir_database* db, const kspir_keys* const* keys,
                              int32_t nq, const uint64_t* queries_dev, uint64_t* responses_dev, void* stream);

/* Batched online answer spanning databases of uneven size: nseg databases
 * (1..8, one N1, each with an int8-MFMA shape), seg_queries[s] consecutive
 * queries against dbs[s], keys[q] per query (for its database's r and N1).
 * One launch per stage serves them all: the baby steps of every query, per
 * chunk of blocks one database pass over every database that has rows there
 * and one giant-step launch over their (query, block) pairs, one launch per
 * packing level over the queries deep enough for it.  queries
 * [sum seg_queries][2][2][N] -> responses [sum seg_queries][2][N]; response i
 * equals kspir_answer on query i alone against its own database.  Needs the
 * default pipeline options (mfma_direct, giant_dma, ext_fused, giant_wgs_per_cu
 * 1).  No counterpart in the reference, which answers one query against one
 * database per call. */
int kspir_answer_multi(kspir_context* ctx, int32_t nseg, const kspir_database* const* dbs,
                       const int32_t* seg_queries, const kspir_keys* const* keys, const uint64_t* queries,
                       uint64_t* responses);
int kspir_answer_multi_device(kspir_context* ctx, int32_t nseg, const kspir_database* const* dbs,
                              const int32_t* seg_queries, const kspir_keys* const* keys,
                              const uint64_t* queries_dev, uint64_t* responses_dev, void* stream);

/* matrix_vector_mul_bsgs_rns_crt_large (twosteps.h:55).  result: [r][2][N]
 * NTT form.  Like the reference, query[0] (a,b mod Q) is replaced by the
 * modulus-switched query in NTT form (twosteps.cpp:1074-1086); query[1] is
 * left as given.  Host pointers. */
int kspir_matrix_vector_mul_bsgs_rns_crt_large(kspir_context* ctx, const kspir_database* db,
                                               const kspir_keys* keys, uint64_t* query,
                                               uint64_t* result);

/* externalProduct (lwe.h:292).  result [2][N] is accumulated into, as NTT-
 * domain words, and returned in coefficient form.  first_dim [2][N] is read
 * in NTT form when first_dim_is_ntt != 0 (and then overwritten with its
 * coefficient form, as the reference does), else in coefficient form. */
int kspir_external_product(kspir_context* ctx, const kspir_keys* keys, uint64_t* result,
                           uint64_t* first_dim, int first_dim_is_ntt);

/* packingRLWEs (answer.h:79).  rlwes [r][2][N] in coefficient form (NTT form
 * when rlwes_are_ntt != 0), r = the keys' r; their contents are unspecified
 * afterwards (the reference also mutates them).  result [2][N] coefficient. */
int kspir_packing_rlwes(kspir_context* ctx, const kspir_keys* keys, uint64_t* result,
                        uint64_t* rlwes, int rlwes_are_ntt);

/* Inspection of one answer (test and profiling aid): optional outputs, each
 * may be NULL.  cipherbuf [N][N1][2] is reorientCipher's buffer
 * (twosteps.cpp:1386), firstdim [r][2][N] the first-dimension results (NTT
 * form), ext [r][2][N] the external-product outputs (coefficient form). */
int kspir_answer_inspect(kspir_context* ctx, const kspir_database* db, const kspir_keys* keys,
                         const uint64_t* query, uint64_t* response, uint64_t* cipherbuf,
                         uint64_t* firstdim, uint64_t* ext);

/* Client recover: decrypt_bsgs_total (twosteps.h:30, twosteps.cpp:388-418)
 * for nq responses on the GPU.  secrets [nq][N] (the secret's coefficients
 * mod crtMod), responses [nq][2][N] (coefficient form, as packingRLWEs
 * returns them; not modified, where the reference turns its argument into
 * NTT form), messages [nq][N]; host pointers.  r^-1 mod crtMod is computed
 * for every power-of-two r (the reference's QInv table has no entry for
 * r = 128 and 1024, where it multiplies by 0). */
int kspir_decrypt_bsgs_total(kspir_context* ctx, int32_t r, int32_t nq, const uint64_t* secrets,
                             const uint64_t* responses, uint64_t* messages);

/* Giant-step half of genAutoKeyFromOffline (twosteps.h, twosteps.cpp:1814-1863,
 * driven by tests/test-two-steps-key.cpp:118-146) for nc clients at once, on
 * the GPU: the chain of evalAutoRNSKey (twosteps.cpp:1588-1715, auxMod
 * branch) by 5^N1 from the auto key at 5^N1, then the modulus switch to
 * crtMod.  Host pointers:
 *   offline2 [nc][10][2][N]  keyMap.at(5^N1) of AutoKeyBSGSRNS(N, crtMod, auxMod, 5, 2^14, 2^14)
 *                            (rows 0..4 mod crtMod, 5..9 mod auxMod, NTT form)
 *   auto2    [nc][6][2][N]   keyMap.at(5^N1) of AutoKeyBSGSRNS(N, crtMod, auxMod, 3, 2^11*auxMod, 2^15)
 *   chain    [nc][N2-2][6][2][N]  the keys the reference inserts into the auto key at 5^(k N1),
 *                            k = 2..N2-1, coefficient form (may be NULL)
 *   giant    [nc][N2-1][3][2][N]  result.keyMap.at(5^(k N1)), k = 1..N2-1, NTT form mod crtMod
 * Each RLWE is (a, b).  N1 is a power of two in [2, N/4].  As in the
 * reference, the key at k = 1 (giant[:, 0]) is switched from auto2's words
 * as stored, in NTT form, by a switch that expects coefficient form: its
 * words equal the reference's, but that key does NOT decrypt and must not be
 * used as a giant-step key (auto2 itself has another gadget, so it is no
 * substitute either): a caller that needs the key at 5^N1 generates it
 * directly (bsgsKeyGen).  Every other giant[:, k] is a valid key.  Device state and pinned
 * output staging are kept on the context between calls.  The crtBaMod
 * (baby-step) half is not provided. */
int kspir_gen_giant_keys_from_offline(kspir_context* ctx, int32_t n1, int32_t nc, const uint64_t* offline2,
                                      const uint64_t* auto2, uint64_t* chain, uint64_t* giant);

/* Client side of test-pir.cpp:20-24,68-89 for nc clients at once, on the
 * GPU.  For client i, as after srand(seeds[i]): the Secret (secret.cpp:13-60),
 * query_bsgs_rns(target_col[i]) (twosteps.h:26, twosteps.cpp:279-318),
 * RGSWCiphertext::keyGen(secret, target_packing[i]) (lwe.h:141,
 * lwe.cpp:297-352), AutoKeyBSGSRNS::bsgsKeyGen (lwe.h:434, lwe.cpp:1143-1160)
 * and AutoKey::keyGen(secret, r) for packing (lwe.h:213, lwe.cpp:509-550),
 * drawing glibc rand() in exactly the reference's order (the generator's
 * stream is produced on the device by jump-ahead).  Host outputs, any may be
 * NULL:  secrets [nc][N] (coefficients mod crtMod), queries [nc][2][2][N],
 * rgsw [nc][4][2][N], baby [nc][N1/2][6][2][N], giant [nc][N2-1][2][2][N],
 * packing [nc][log2 r][4][2][N]: the layouts kspir_keys_create and
 * kspir_answer take.  The query message rounding reproduces the
 * reference's long double round(m * crtMod / p). */
int kspir_client_generate(kspir_context* ctx, int32_t r, int32_t n1, int32_t nc, const uint32_t* seeds,
                          const int32_t* target_col, const int32_t* target_packing, uint64_t* secrets,
                          uint64_t* queries, uint64_t* rgsw, uint64_t* baby, uint64_t* giant, uint64_t* packing);
/* rand() draws one client of shape (r, N1) consumes after srand (host-only) */
int kspir_client_draws(int32_t r, int32_t n1, uint64_t* draws);

/* The same client operations one at a time, on an explicit secret and on
 * rand() outputs the caller drew in the reference's order (the C++ drop-ins
 * of kspir_pir.h call rand() themselves, so the caller's global generator
 * advances exactly as with the reference).  secret: [N] coefficients mod
 * crtMod (0, 1 or crtMod-1); draws: ndraws rand() outputs; out (host):
 *   KSPIR_CLIENT_QUERY   query_bsgs_rns(row = param): 5N draws, [2][2][N]
 *   KSPIR_CLIENT_RGSW    RGSWCiphertext::keyGen(index = param): 12N draws, [4][2][N]
 *   KSPIR_CLIENT_BSGS    bsgsKeyGen(N1 = n1): 15N*N1/2 + 6N*(N2-1) draws,
 *                        baby [N1/2][6][2][N] then giant [N2-1][2][2][N]
 *   KSPIR_CLIENT_PACKING AutoKey::keyGen(num = param, packing_rlwe = true):
 *                        indices N/i+1 for i = 1, 2, 4, .. < num; 12N draws and
 *                        [4][2][N] per index
 *   KSPIR_CLIENT_RGSW_FORWARD    keyGen(index, is_reverse = false): message X^index
 *   KSPIR_CLIENT_PACKING_LWE     keyGen(num, packing_rlwe = false): indices
 *                        i+1 for i = 2, 4, .. <= num+1 (lwe.cpp:509-550) */
#define KSPIR_CLIENT_QUERY 0
#define KSPIR_CLIENT_RGSW 1
#define KSPIR_CLIENT_BSGS 2
#define KSPIR_CLIENT_PACKING 3
#define KSPIR_CLIENT_RGSW_FORWARD 4
#define KSPIR_CLIENT_PACKING_LWE 5
int kspir_client_encrypt(kspir_context* ctx, int32_t kind, const uint64_t* secret, int64_t param, int32_t n1,
                         const uint32_t* draws, uint64_t ndraws, uint64_t* out);

/* Serving loop (no counterpart in the reference, whose driver answers one
 * query at a time).  A worker thread answers submitted queries in batches:
 * whenever the GPU is free it takes the database of the oldest waiting query
 * and up to max_batch of that database's waiting queries, in arrival order,
 * as many as workspace_budget_bytes admits at the per-query device footprint
 * of the database's shape (0 = half of the device's free memory at
 * creation).  With the context option "serve_multi" set to 1 (0 by default)
 * and the options of kspir_answer_multi, waiting queries of other databases of the
 * same N1 and at most as many blocks fill the rest of the batch in arrival
 * order (at most 8 databases), each admitted while the whole batch fits the
 * budget at the first database's per-query footprint, and the batch is
 * answered as by kspir_answer_multi: one launch per stage.  A larger
 * database never joins: its longer giant steps would delay every query of
 * the batch.  The server drives the context:
 * while it exists, do not call other functions on that context.  query ([2][2][N], host) and response
 * ([2][N], host) must stay valid until kspir_server_wait returns for the
 * ticket; latency_ms is submit-to-completion time.  Responses equal
 * kspir_answer's for the same query. */
typedef struct kspir_server kspir_server;
int kspir_server_create(kspir_context* ctx, int32_t max_batch, size_t workspace_budget_bytes, kspir_server** out);
/* Same with `lanes` (1..8) worker threads: lane 0 drives ctx, each other
 * lane a context of its own on ctx's device (with ctx's options), each
 * taking the next batch as soon as it is free, so that batches overlap on
 * the device (one batch's database stream beside another's transforms).
 * The budget is split evenly between the lanes. */
int kspir_server_create_lanes(kspir_context* ctx, int32_t max_batch, size_t workspace_budget_bytes, int32_t lanes,
                              kspir_server** out);
/* A query whose batch of one needs more device workspace than a lane's
 * budget (kspir_batch_footprint_bytes with nq = 1) is rejected here with
 * KSPIR_EBUDGET: no batch could ever hold it. */
int kspir_server_submit(kspir_server* server, const kspir_database* db, const kspir_keys* keys,
                        const uint64_t* query, uint64_t* response, uint64_t* ticket);
/* Returns the request's own result: KSPIR_ECANCELED after kspir_server_cancel,
 * KSPIR_EINJECTED for a batch failed on purpose (serve_fail_batches), or the
 * error of its batch's answer. */
int kspir_server_wait(kspir_server* server, uint64_t ticket, double* latency_ms);
/* Withdraws a request that no batch has taken yet: its wait then returns
 * KSPIR_ECANCELED and its response buffer is left untouched.  KSPIR_EBUSY
 * when a batch already holds it (it completes normally), KSPIR_EINVAL for a
 * ticket never issued or already collected. */
int kspir_server_cancel(kspir_server* server, uint64_t ticket);
int kspir_server_stats(kspir_server* server, uint64_t* batches, uint64_t* answered, uint64_t* pending);
/* batches so far that spanned more than one database */
int kspir_server_spanning_batches(kspir_server* server, uint64_t* spanning);
/* waits for every submitted query, then stops the worker */
void kspir_server_destroy(kspir_server* server);

#ifdef __cplusplus
}
#endif
#endif
