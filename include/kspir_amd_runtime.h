/*
 * kspir_amd_runtime.h — device-memory, timing and test plumbing of the
 * MI355X KsPIR library.  These entry points have no counterpart in the
 * reference (which is host-only); they exist so that callers without a GPU
 * framework (the ctypes tests, the benchmark) can keep inputs resident in HBM
 * and time kernels with HIP events on the context's stream.
 */
#ifndef KSPIR_AMD_RUNTIME_H
#define KSPIR_AMD_RUNTIME_H

#include <stdint.h>
#include <stddef.h>
#include "kspir_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

int kspir_device_count(int* count);

/* Tuning options of a context:
 *   "mfma_min_batch"  batches of at least this many queries take the int8-MFMA
 *                     database pass (default 2); smaller ones the VALU pass
 *   "mfma_max_chunk"  most queries per MFMA database sweep (default: as many
 *                     as the LDS holds)
 *   "mfma_direct"     1 (default): one kernel writes the MFMA products in the
 *                     first-dimension layout directly (16 slots per
 *                     workgroup, query limbs streamed from L2 into an LDS ring
 *                     by DMA, several column tiles ahead); 0: the MFMA pass
 *                     writes tile-major products that a second kernel moves
 *                     into that layout (1 GiB more HBM traffic and scratch per
 *                     64-query step at r 16)
 *   "ext_fused"       1 (default): with the DMA giant-step kernel, the external
 *                     products run in the same workgroups (no coefficient
 *                     planes written out); 0: separate kernels
 *   "chunk_blocks"    batches on that path run the database pass, giant steps
 *                     and external products this many blocks at a time (a
 *                     power of two, default 32; 0: the whole database), so
 *                     their first-dimension buffers hold one chunk of blocks
 *   "fused_min_batch" batches of at least this many queries run the giant
 *                     steps and first-dimension sum as one workgroup per
 *                     (query, block) (default 16); smaller ones spread them
 *                     over one workgroup per (block, row, prime)
 *   "fused_pack"      1 (default): for batches of at least fused_min_batch
 *                     queries each packing level is one kernel with a
 *                     workgroup per (pair, query); 0: three kernels per level
 *   "graphs"          1 (default): kspir_answer_device / _batch_device replay
 *                     a captured hipGraph of the whole pipeline for buffers
 *                     and batch sizes seen before (captured after a first,
 *                     eager call); 0: launch kernel by kernel
 *   "pack_fused_min_wgs" with fused_pack, a packing level runs fused only when
 *                     it has at least this many (pair, query) workgroups
 *                     (default 0: every level)
 *   "giant_wgs_per_cu" 1 (default): the fused giant-step kernel with six LDS
 *                     planes, one workgroup per CU; 2: a variant in four
 *                     planes and 128 VGPRs (b-rows gathered from L2), two per CU
 *   "giant_dma"       with giant_wgs_per_cu 1: 1 (default) takes the fused
 *                     giant-step kernel that prefetches each next row into
 *                     LDS by DMA while the current row is transformed (eight
 *                     planes); 0: rows loaded through registers (six planes)
 *   "db_waves"        workgroup shape of the single-query database pass:
 *                     2, 4 (default), 8 or 16 waves on 64 slots, each wave
 *                     taking N2 / waves rows
 *   "serve_multi"     1: a serving loop on this context fills a batch with
 *                     waiting queries of several databases of one N1
 *                     (kspir_server_create); 0 (default): one database per
 *                     batch, which measured lower latency at equal
 *                     throughput on mixed traffic
 *   "serve_graphs"    0 (default): a serving loop launches each batch
 *                     eagerly (its batch sizes vary, so most batches would be
 *                     a new capture); 1: it captures and replays graphs as
 *                     the answer calls above do
 *   "serve_fail_batches"  n >= 0 (default 0): the next n batches a serving
 *                     loop runs on this context fail with KSPIR_EINJECTED,
 *                     launching nothing (for tests of the failure path);
 *                     set it before kspir_server_create
 * Setting an option waits for the context stream and drops captured graphs. */
int kspir_context_set(kspir_context* ctx, const char* key, int64_t value);
int kspir_context_get(kspir_context* ctx, const char* key, int64_t* value);

/* Autotuner (no counterpart in the reference, whose shapes are constants of
 * a CPU loop).  Times the answer pipeline for this database on this device,
 * with the given client keys reused in turn to fill batches of up to
 * max_batch queries, and sets the options above: db_waves (fastest database
 * pass), mfma_min_batch and fused_min_batch (smallest batch size from which
 * that path stays faster, over powers of two up to max_batch),
 * mfma_max_chunk, giant_wgs_per_cu, giant_dma, fused_pack and pack_fused_min_wgs
 * (fastest at max_batch).  *serve_batch (may be
 * NULL) receives the batch size with the most queries/s whose step time is
 * within latency_budget_ms (0: no bound).  report (may be NULL): every
 * timing as JSON text, cut to report_len bytes.  Runs for a few seconds. */
int kspir_autotune(kspir_context* ctx, const kspir_database* db, const kspir_keys* const* keys, int32_t nkeys,
                   int32_t max_batch, double latency_budget_ms, int32_t* serve_batch, char* report,
                   size_t report_len);

/* device bytes per query of a 64-query batch of shape (r, N1) under the
 * default options (workspace and batched database-pass scratch, divided by
 * 64).  The serving loop admits each batch by its exact footprint for the
 * batch size and the lane's options. */
int kspir_query_footprint_bytes(int32_t r, int32_t n1, size_t* bytes);
/* device bytes of one batch of nq queries of shape (r, N1) under ctx's
 * options (NULL: the defaults): the figure the serving loop admits a batch
 * by.  It is not monotonic in nq: a few queries of a large database can take
 * the unchunked path, whose buffers hold every block's rows per query. */
int kspir_batch_footprint_bytes(const kspir_context* ctx, int32_t r, int32_t n1, int32_t nq, size_t* bytes);
int kspir_device_alloc(kspir_context* ctx, size_t bytes, void** out);
int kspir_device_free(kspir_context* ctx, void* ptr);
int kspir_memcpy_h2d(kspir_context* ctx, void* dst, const void* src, size_t bytes);
int kspir_memcpy_d2h(kspir_context* ctx, void* dst, const void* src, size_t bytes);
int kspir_synchronize(kspir_context* ctx);
/* hipStream_t of the context */
void* kspir_context_stream(kspir_context* ctx);

/* HIP-event timing on the context stream: `warmup` untimed then `iters`
 * timed kspir_answer_device calls; *ms = total milliseconds of the timed
 * calls. */
int kspir_time_answer(kspir_context* ctx, const kspir_database* db, const kspir_keys* keys,
                      const uint64_t* query_dev, uint64_t* response_dev, int warmup, int iters,
                      float* ms);

/* same for a batch of nq queries (kspir_answer_batch_device per call) */
int kspir_time_answer_batch(kspir_context* ctx, const kspir_database* db, const kspir_keys* const* keys,
                            int32_t nq, const uint64_t* queries_dev, uint64_t* responses_dev, int warmup,
                            int iters, float* ms);

/* same for a batch spanning databases (kspir_answer_multi_device per call) */
int kspir_time_answer_multi(kspir_context* ctx, int32_t nseg, const kspir_database* const* dbs,
                            const int32_t* seg_queries, const kspir_keys* const* keys, const uint64_t* queries_dev,
                            uint64_t* responses_dev, int warmup, int iters, float* ms);

/* HIP-event timing of the database-pass kernel alone (after one answer has
 * filled the rotation words): *ms_per_launch = average milliseconds. */
int kspir_time_db_pass(kspir_context* ctx, const kspir_database* db, const kspir_keys* keys,
                       const uint64_t* query_dev, int iters, float* ms_per_launch);

/* HIP-event timing of the batched (int8-MFMA) database pass of nq >= 2
 * queries, one kernel at a time (after one first dimension has filled the
 * rotation words): ms3[0] the query limbs (k_ct_limbs), ms3[1] the MFMA
 * products (with mfma_direct: k_db_mfma_direct writing the slot-group-major
 * first-dimension layout of the batched pipeline; otherwise k_db_mfma),
 * ms3[2] their move into the first-dimension layout (k_mfma_scatter; no work
 * with mfma_direct), each the average milliseconds per launch over iters. */
int kspir_time_db_pass_batch(kspir_context* ctx, const kspir_database* db, const kspir_keys* const* keys,
                             int32_t nq, const uint64_t* queries_dev, int iters, float* ms3);

/* Plain device-to-device copy kernel (16-byte loads and stores, `bytes` a
 * multiple of 16) timed with HIP events: *ms_per_launch = average
 * milliseconds.  It is the reference kernel for HBM counters (FETCH_SIZE /
 * WRITE_SIZE of a copy of known size) and for achievable copy bandwidth. */
int kspir_time_copy(kspir_context* ctx, void* dst_dev, const void* src_dev, size_t bytes, int iters,
                    float* ms_per_launch);

/* Test hook: `count` polynomials of N u32 residues in prime p (0: q1,
 * 1: q2, 2: bsMod, 3: plaintext p, 4: auxMod), forward (inverse != 0:
 * inverse) NTT in place, host memory. */
int kspir_test_ntt(kspir_context* ctx, uint32_t* polys, int count, int prime, int inverse);

#ifdef __cplusplus
}
#endif
#endif
