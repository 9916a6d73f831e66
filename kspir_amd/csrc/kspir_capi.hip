// kspir_capi.hip — host side of the MI355X KsPIR library: device tables,
// resident database and key handles, and the launch sequence of the online
// answer behind the C ABI of include/kspir_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "kspir_kernels.h"
#include "kspir_client.h"
#include "kspir_offline.h"
#include "kspir_rng.h"
#include "../../include/kspir_amd.h"
#include "../../include/kspir_amd_runtime.h"

using namespace kspir;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
int fail_hip(const char* what, hipError_t e) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? KSPIR_ENOMEM : KSPIR_EHIP;
}

#define HIP_OK(x)                                        \
    do {                                                 \
        hipError_t e_ = (x);                             \
        if (e_ != hipSuccess) return fail_hip(#x, e_);   \
    } while (0)

#define LAUNCH_OK(name)                                  \
    do {                                                 \
        hipError_t e_ = hipGetLastError();               \
        if (e_ != hipSuccess) return fail_hip(name, e_); \
    } while (0)

// ------------------------------------------------------------ host arithmetic

using u64 = uint64_t;
using u128 = unsigned __int128;

u64 powmod(u64 b, u64 e, u64 q) {
    u64 r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = (u64)((u128)r * b % q);
        b = (u64)((u128)b * b % q);
        e >>= 1;
    }
    return r;
}
u64 invmod(u64 a, u64 q) { return powmod(a, q - 2, q); }   // q prime
uint32_t bitrev12(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < kLogN; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}
uint32_t shoup(uint32_t w, uint32_t q) { return (uint32_t)(((u64)w << 32) / q); }

// smallest primitive 2N-th root of unity mod a prime q (x^N == -1)
u64 min_root(u64 q) {
    for (u64 x = 2; x < q; ++x)
        if (powmod(x, kN, q) == q - 1) return x;
    return 0;
}

bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
int ilog2(int v) { int b = 0; while (v > 1) { v >>= 1; ++b; } return b; }

}  // namespace

// --------------------------------------------- no-allocation answer scratch
// Workspace of kspir_answer_device_noalloc: the buffers of one single-query
// answer, up to the 256 MB test-pir shape (r 16, N1 128 needs 61.6 MB).  It
// is a file-scope device array, so every device the library runs on holds
// its own copy until the process ends.  In host code `scratch` only names
// the symbol: its address on each device comes from hipGetSymbolAddress at
// context creation (with that device current), and host code passes only
// that per-device pointer on.
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) uint32_t scratch[kScratchBytes / 4];
static std::mutex g_scratch_mu;
static uint32_t* g_scratch_ptr[kMaxDevices] = {};

// ------------------------------------------------------------------ handles

struct kspir_context {
    int device = 0;
    hipStream_t stream = nullptr;
    Tabs T{};
    std::vector<void*> tables;
    // per-call workspace: one allocation, carved for (r, N1, nq) on each call
    void* ws = nullptr;
    size_t ws_bytes = 0;
    QS S{};
    uint64_t* q_dev = nullptr;      // [nq][4][N]
    uint64_t* resp_dev = nullptr;   // [nq][2][N]
    uint32_t *D0, *D1, *bhat0, *mid, *icoef, *ctw, *dbout, *gcoef, *gks, *fd, *ecoef, *epart;
    uint32_t *extA, *extB, *pcoef, *ppart;
    uint64_t* u64tmp = nullptr;     // [max(r,2)][2][N]
    uint64_t* u64tmp2 = nullptr;    // [N][N1][2] packed rotation words
    // per-batch key pointer table
    KeyPtrs* kp_dev = nullptr;
    KeyPtrs* kp_host = nullptr;     // pinned staging
    int kp_cap = 0, kp_count = 0;
    // batched database pass scratch (query limbs + tile-major products)
    void* mws = nullptr;
    size_t mws_bytes = 0;
    int mfma_min_batch = 2;         // batches of at least this many queries use the MFMA pass
    int mfma_max_chunk = 1 << 30;   // cap on queries per MFMA database sweep (LDS also caps it)
    int ext_fused = 1;              // the DMA giant-step kernel also runs the external products
    int chunk_blocks = 32;          // blocks per chunk of the batched pipeline (0: the whole database at once)
    int cur_cb = 0;                 // chunk size the current workspace carve was made for (0: not chunked)
    int mfma_direct = 1;            // 1: MFMA products written straight into dbout (k_db_mfma_direct); 0: tile-major + scatter
    int fused_min_batch = 16;       // batches of at least this many queries use k_giant_fused
    int fused_pack = 1;             // one k_pack_level per packing level (0: three kernels per level)
    int db_waves = 4;               // waves per workgroup of the single-query database pass (2, 4, 8, 16)
    int giant_wgs_per_cu = 1;       // k_giant_fused (1) or its two-per-CU variant k_giant_fused2 (2)
    int giant_dma = 1;              // one per CU with rows prefetched by LDS-DMA (k_giant_fused_dma)
    int64_t pack_fused_min_wgs = 0; // a packing level is fused only with at least this many (pair, query) workgroups
    int serve_multi = 0;            // the serving loop puts queries of several databases in one batch (one launch per stage)
    int serve_graphs = 0;           // the serving loop replays captured graphs (1) or launches eagerly (0)
    int64_t serve_fail_batches = 0; // the serving loop's next batches that fail on purpose (KSPIR_EINJECTED)
    // captured pipelines: one hipGraph per (database, key table, batch size,
    // query / response buffers, stream); dropped whenever a buffer they name
    // is reallocated or freed, or an option changes
    struct Graph {
        const void* db;
        const void* kp;
        int nq;
        const void* q;
        void* resp;
        hipStream_t st;
        hipGraphExec_t exec;   // null: capture failed, run eagerly
        std::vector<const void*> segs;   // a batch spanning databases: (database, query count) per segment
    };
    std::vector<Graph> graphs;
    std::mutex graph_mu;   // the cache is also dropped from threads that free handles
    hipStream_t last_stream = nullptr;   // last caller stream work was enqueued on (besides `stream`)
    int use_graphs = 1;
    // client generation: M^kRngChunk of the rand() window, and the query
    // encoding round(m * Q / p) in long double for every m < p
    uint32_t* rng_jump = nullptr;
    uint64_t* enc_tab = nullptr;
    // offline key derivation: device state and pinned output staging, grown on demand
    void* ows = nullptr;
    size_t ows_bytes = 0;
    void* ohost = nullptr;
    size_t ohost_bytes = 0;
};

static void drop_graphs(kspir_context* c, const void* only = nullptr) {
    std::lock_guard<std::mutex> lock(c->graph_mu);
    std::vector<kspir_context::Graph> keep;
    for (auto& g : c->graphs) {
        const bool names = g.db == only || g.kp == only || std::find(g.segs.begin(), g.segs.end(), only) != g.segs.end();
        if (only && !names) { keep.push_back(g); continue; }
        if (g.exec) hipGraphExecDestroy(g.exec);
    }
    c->graphs.swap(keep);
}

// every live context, so that freeing a database or key handle drops the
// graphs that name it in all contexts of its device (any context of the
// device may use a handle)
static std::mutex g_reg_mu;
static std::vector<kspir_context*> g_reg;

static void drop_graphs_everywhere(int device, const void* only) {
    std::lock_guard<std::mutex> lock(g_reg_mu);
    for (kspir_context* c : g_reg)
        if (c->device == device) {
            hipStreamSynchronize(c->stream);
            drop_graphs(c, only);
        }
}

struct kspir_database {
    kspir_context* ctx;
    int r, n1, n2;
    uint64_t* words;            // [r][N2][N1][N]
    int8_t* limbs = nullptr;    // [N][2][4][r*N2][N1] 7-bit limbs, built on the first batched call
    std::mutex limbs_mu;        // contexts answering batches at once build them once
};

struct kspir_keys {
    kspir_context* ctx;
    int r, n1, n2;
    KeyPtrs ptrs{};
    KeyPtrs* dev_ptrs = nullptr;   // one-entry device copy of ptrs
    uint32_t* baby = nullptr;
    uint32_t* giant = nullptr;
    uint32_t* packing = nullptr;
    uint32_t* rgsw = nullptr;
    uint16_t* perms = nullptr;
};

namespace {

int build_tables(kspir_context* c) {
    const uint32_t qs[5] = {kQ1, kQ2, kBs, kP, kAux};
    const u64 root_q = 3375402822066082ULL;   // crt.h:81 root_of_unity_crt
    const u64 roots[5] = {root_q % kQ1, root_q % kQ2, min_root(kBs), min_root(kP), min_root(kAux)};
    for (int p = 0; p < 5; ++p) {
        const uint32_t q = qs[p];
        const u64 w = roots[p];
        std::vector<uint32_t> fw(kN), fwsh(kN), iv(kN), ivsh(kN), pw(2 * kN);
        u64 x = 1;
        const u64 winv = invmod(w, q);
        u64 y = 1;
        for (int i = 0; i < kN; ++i) {
            const uint32_t k = bitrev12(i);
            fw[k] = (uint32_t)x;
            iv[k] = (uint32_t)y;
            x = (u64)((u128)x * w % q);
            y = (u64)((u128)y * winv % q);
        }
        for (int i = 0; i < kN; ++i) { fwsh[i] = shoup(fw[i], q); ivsh[i] = shoup(iv[i], q); }
        u64 e = 1;
        for (int i = 0; i < 2 * kN; ++i) {
            pw[i] = (uint32_t)(((u128)e << 32) % q);   // Montgomery form
            e = (u64)((u128)e * w % q);
        }
        std::vector<uint32_t> pairs(4 * kN);
        for (int i = 0; i < kN; ++i) {
            pairs[2 * i] = 0u - fw[i]; pairs[2 * i + 1] = fwsh[i];   // forward: 2^32 - w (see ntt_fwd_pass)
            pairs[2 * kN + 2 * i] = iv[i]; pairs[2 * kN + 2 * i + 1] = ivsh[i];
        }
        uint32_t* d = nullptr;
        HIP_OK(hipMalloc(&d, sizeof(uint32_t) * (4 * kN + 2 * kN + 4 * kN)));
        c->tables.push_back(d);
        HIP_OK(hipMemcpy(d + 0 * kN, fw.data(), 4 * kN, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d + 1 * kN, fwsh.data(), 4 * kN, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d + 2 * kN, iv.data(), 4 * kN, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d + 3 * kN, ivsh.data(), 4 * kN, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d + 4 * kN, pw.data(), 8 * kN, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d + 6 * kN, pairs.data(), 16 * kN, hipMemcpyHostToDevice));
        PrimeTab& P = c->T.p[p];
        P.q = q;
        uint32_t inv = 1;   // q^-1 mod 2^32 by Newton iteration
        for (int it = 0; it < 5; ++it) inv *= 2u - q * inv;
        P.mont_qinv = (uint32_t)(0u - inv);
        P.ninv = (uint32_t)invmod(kN, q);
        P.ninv_sh = shoup(P.ninv, q);
        P.fw = d;
        P.fw_sh = d + kN;
        P.iv = d + 2 * kN;
        P.iv_sh = d + 3 * kN;
        P.psi_pow = d + 4 * kN;
        P.fw2 = reinterpret_cast<const uint2*>(d + 6 * kN);
        P.iv2 = reinterpret_cast<const uint2*>(d + 8 * kN);
    }
    std::vector<uint16_t> se(kN), es(kN);
    for (int s = 0; s < kN; ++s) {
        const uint32_t ex = 2 * bitrev12(s) + 1;
        se[s] = (uint16_t)ex;
        es[ex >> 1] = (uint16_t)s;
    }
    uint16_t* t16 = nullptr;
    HIP_OK(hipMalloc(&t16, sizeof(uint16_t) * 2 * kN));
    c->tables.push_back(t16);
    HIP_OK(hipMemcpy(t16, se.data(), 2 * kN, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(t16 + kN, es.data(), 2 * kN, hipMemcpyHostToDevice));
    Tabs& T = c->T;
    T.slot_exp = t16;
    T.exp_slot = t16 + kN;
    T.q1inv_mod_q2 = (uint32_t)invmod(kQ1 % kQ2, kQ2);
    T.q1inv_mod_q2_sh = shoup(T.q1inv_mod_q2, kQ2);
    for (int p = 0; p < 2; ++p) {
        T.bsinv_q[p] = (uint32_t)invmod(kBs, qs[p]);
        T.bsinv_q_sh[p] = shoup(T.bsinv_q[p], qs[p]);
    }
    T.q_mod_bs = (uint32_t)(kQ % kBs);
    T.qinv_bs = (uint32_t)invmod(T.q_mod_bs, kBs);
    T.qinv_bs_sh = shoup(T.qinv_bs, kBs);
    return KSPIR_OK;
}


// (query, block) pairs that keep the GPU full with one fused giant-step
// workgroup each (eight rounds of one per CU)
constexpr int64_t kFusedFillWgs = 2048;

// whether a batch of nq queries over r blocks takes the fused giant-step
// kernel: from fused_min_batch queries, whenever its (query, block) pairs
// fill the GPU, and from 4 queries at 64 blocks or more.  The unfused
// kernels keep every block's giant rows (gcoef, gks) per query, 2.6 GiB at
// r = 1024, so a large database takes the fused path even for a few queries
// (r 1024 from 2).  Measured at 1 GB (r 64): 4 queries 2.02 ms fused against
// 2.09 unfused, 8 queries 2.17 against 2.71; at 4 GB (r 256): 4 queries 7.25
// against 7.10 ms, at a ninth of the workspace; 2 and 3 queries stay
// unfused there (5.8 against 7.0 ms for 2).
bool fused_batch(const kspir_context* c, int r, int nq) {
    return nq >= c->fused_min_batch || (int64_t)r * nq >= kFusedFillWgs || (r >= 64 && nq >= 4);
}

// Blocks per chunk of the batched pipeline for nq queries of shape (r, n1),
// or 0 when that batch takes the unchunked path.  A chunked batch runs the
// MFMA database pass, the giant steps and the external products one chunk of
// blocks at a time, so its first-dimension buffers hold one chunk; only the
// int8-MFMA pass with the DMA giant-step kernel and fused external products
// is chunked.
int chunk_blocks_for(const kspir_context* c, int r, int n1, int nq) {
    const int n2 = kN / 2 / n1;
    const bool mfma = n1 % 64 == 0 && n1 <= 256 && (r * n2) % 16 == 0;
    if (!mfma || nq < c->mfma_min_batch || !fused_batch(c, r, nq) || !c->giant_dma || c->giant_wgs_per_cu != 1 ||
        !c->ext_fused)
        return 0;
    int cb = c->chunk_blocks > 0 && c->chunk_blocks < r ? c->chunk_blocks : r;
    while ((cb * n2) % 16) cb *= 2;   // whole 16-row MFMA tiles per chunk
    return cb > r ? r : cb;
}

size_t layout_workspace(kspir_context* c, int r, int n1, int nq, uint32_t* base, bool answer_only, int cb = 0);
int mfma_sweep_queries(const kspir_context* c, int n1);

// Device bytes a batch of nq queries of shape (r, n1) needs beyond the
// database itself: the carved workspace and the MFMA pass's scratch (query
// limbs; tile-major products of one chunk's rows).  The serving loop admits
// batches against it.
size_t batch_footprint_bytes(kspir_context* c, int r, int n1, int nq) {
    const int cb = chunk_blocks_for(c, r, n1, nq);
    size_t bytes = layout_workspace(c, r, n1, nq, nullptr, false, cb) * 4;
    const int n2 = kN / 2 / n1;
    if (n1 % 64 == 0 && n1 <= 256 && (r * n2) % 16 == 0 && nq >= c->mfma_min_batch) {
        const int nqc = nq < mfma_sweep_queries(c, n1) ? nq : mfma_sweep_queries(c, n1);
        const size_t C = (size_t)(2 * nqc + 15) / 16 * 16;
        const size_t rows = (size_t)(cb ? cb : r) * n2;
        bytes += (size_t)kN * 2 * 4 * C * n1 + (c->mfma_direct ? 0 : 2 * rows * C * kN * 4);
    }
    return bytes;
}

// Layout of the answer workspace for nq queries of shape (r, n1): sets c->S,
// and every buffer pointer of the context when base is not null; returns
// the size in 32-bit words.  answer_only leaves out the host-transfer and
// inspection areas (q_dev, resp_dev, u64tmp, u64tmp2), which the answer
// kernels themselves never touch.
size_t layout_workspace(kspir_context* c, int r, int n1, int nq, uint32_t* base, bool answer_only, int cb) {
    const int n2 = kN / 2 / n1;
    const size_t n = kN;
    const size_t rh = r / 2 > 0 ? r / 2 : 1;
    const size_t R2 = r > 2 ? r : 2;
    QS S;
    S.D = 9 * n; S.bhat = 3 * n; S.mid = 6 * n; S.icoef = 3 * n;
    S.ctw = (size_t)n1 * 4 * n;
    S.dbout = (size_t)r * n2 * 4 * n; S.gcoef = (size_t)r * n2 * 2 * n; S.gks = (size_t)r * n2 * 4 * n;
    S.fd = (size_t)r * 4 * n; S.ecoef = (size_t)r * 4 * n; S.epart = (size_t)r * 16 * n;
    S.ext = (size_t)r * 4 * n; S.pcoef = rh * 2 * n; S.ppart = rh * 16 * n;
    S.qin64 = 4 * n; S.resp64 = 2 * n;
    if (cb > 0) {
        // chunked batch: first-dimension products of one chunk of blocks;
        // the per-row giant buffers, the coefficient planes and the external
        // products' partial sums are not used (fused kernel)
        S.dbout = (size_t)cb * n2 * 4 * n;
        S.gcoef = S.gks = S.fd = S.ecoef = S.epart = 0;
        // every packing level fused (k_pack_level): no per-level scratch
        if (c->fused_pack && c->pack_fused_min_wgs <= nq) S.pcoef = S.ppart = 0;
    }
    auto up = [](size_t w) { return (w + 63) / 64 * 64; };
    struct Part { size_t words; uint32_t** dst; };
    const size_t Q = (size_t)nq;
    std::vector<Part> parts = {
        {Q * S.D, &c->D0}, {Q * S.D, &c->D1}, {Q * S.bhat, &c->bhat0}, {Q * S.mid, &c->mid},
        {Q * S.icoef, &c->icoef}, {Q * S.ctw, &c->ctw}, {Q * S.dbout, &c->dbout}, {Q * S.gcoef, &c->gcoef},
        {Q * S.gks, &c->gks}, {Q * S.fd, &c->fd}, {Q * S.ecoef, &c->ecoef}, {Q * S.epart, &c->epart},
        {Q * S.ext, &c->extA}, {Q * S.ext, &c->extB}, {Q * S.pcoef, &c->pcoef}, {Q * S.ppart, &c->ppart},
    };
    if (!answer_only) {
        parts.push_back({Q * 2 * S.qin64, (uint32_t**)&c->q_dev});
        parts.push_back({Q * 2 * S.resp64, (uint32_t**)&c->resp_dev});
        parts.push_back({R2 * 4 * n, (uint32_t**)&c->u64tmp});
        parts.push_back({(size_t)n1 * 4 * n, (uint32_t**)&c->u64tmp2});
    }
    size_t total = 0;
    for (auto& p : parts) {
        if (base) *p.dst = base + total;
        total += up(p.words);
    }
    if (base) {
        if (answer_only) { c->q_dev = nullptr; c->resp_dev = nullptr; c->u64tmp = nullptr; c->u64tmp2 = nullptr; }
        c->S = S;
    }
    return total;
}

// carve the workspace for nq queries of shape (r, n1); grows on demand.
// allow_chunk = false keeps whole-database buffers (callers that run the
// first dimension alone).
int carve_workspace_cb(kspir_context* c, int r, int n1, int nq, int cb) {
    const size_t total = layout_workspace(c, r, n1, nq, nullptr, false, cb);
    if (total * 4 > c->ws_bytes) {
        if (c->ws) {
            HIP_OK(hipStreamSynchronize(c->stream));
            if (c->last_stream) HIP_OK(hipStreamSynchronize(c->last_stream));
            drop_graphs(c);
            HIP_OK(hipFree(c->ws));
            c->ws = nullptr;
            c->ws_bytes = 0;
        }
        HIP_OK(hipMalloc(&c->ws, total * 4));
        c->ws_bytes = total * 4;
    }
    layout_workspace(c, r, n1, nq, (uint32_t*)c->ws, false, cb);
    c->cur_cb = cb;
    return KSPIR_OK;
}

int carve_workspace(kspir_context* c, int r, int n1, int nq, bool allow_chunk = true) {
    return carve_workspace_cb(c, r, n1, nq, allow_chunk ? chunk_blocks_for(c, r, n1, nq) : 0);
}

int grid1d(size_t total) {
    size_t g = (total + 255) / 256;
    return (int)(g > 65535 ? 65535 : (g ? g : 1));
}

// workgroup shape of the single-query database pass: WAVES waves of RPW
// rows each (RPW = N2 / WAVES, at least 1); waves = 0 picks 4
template <int WAVES>
int launch_db_product_w(kspir_context* c, const kspir_database* db, int nq, hipStream_t st) {
    const int n1 = db->n1, n2 = db->n2, r = db->r;
    const size_t bw = (size_t)n2 * n1 * kN, obw = (size_t)n2 * 4 * kN;
    const dim3 dg(kN / 64, r, nq), bd(64 * WAVES);
    const QS& S = c->S;
    const Tabs& T = c->T;
    const int rpw = n2 <= WAVES ? 1 : n2 / WAVES;
    if (n2 > WAVES && n2 % WAVES) return fail(KSPIR_EINVAL, "N2 must be a power of two");
    switch (rpw) {
        case 1: hipLaunchKernelGGL((k_db_product<1, WAVES>), dg, bd, 0, st, db->words, c->ctw, c->dbout, n1, n2, bw, obw, S, T); break;
        case 2: hipLaunchKernelGGL((k_db_product<2, WAVES>), dg, bd, 0, st, db->words, c->ctw, c->dbout, n1, n2, bw, obw, S, T); break;
        case 4: hipLaunchKernelGGL((k_db_product<4, WAVES>), dg, bd, 0, st, db->words, c->ctw, c->dbout, n1, n2, bw, obw, S, T); break;
        case 8: hipLaunchKernelGGL((k_db_product<8, WAVES>), dg, bd, 0, st, db->words, c->ctw, c->dbout, n1, n2, bw, obw, S, T); break;
        case 16: hipLaunchKernelGGL((k_db_product<16, WAVES>), dg, bd, 0, st, db->words, c->ctw, c->dbout, n1, n2, bw, obw, S, T); break;
        case 32: hipLaunchKernelGGL((k_db_product<32, WAVES>), dg, bd, 0, st, db->words, c->ctw, c->dbout, n1, n2, bw, obw, S, T); break;
        default: return fail(KSPIR_EINVAL, "unsupported N2 for the database pass (rows per wave above 32)");
    }
    LAUNCH_OK("k_db_product");
    return KSPIR_OK;
}

int launch_db_product(kspir_context* c, const kspir_database* db, int nq, hipStream_t st) {
    switch (c->db_waves) {
        case 2: return launch_db_product_w<2>(c, db, nq, st);
        case 8: return launch_db_product_w<8>(c, db, nq, st);
        case 16: return launch_db_product_w<16>(c, db, nq, st);
        default: return launch_db_product_w<4>(c, db, nq, st);
    }
}

bool mfma_shape(const kspir_database* db) {
    return db->n1 % 64 == 0 && db->n1 <= 256 && (db->r * db->n2) % 16 == 0;
}

// batched database pass on int8 MFMA (one database read per chunk of up to
// nqc_max queries); results land in c->dbout exactly as k_db_product's
// parts (timing aid): bit 0 the query limbs, bit 1 the MFMA products, bit 2
// the move into dbout; the whole pass is 7
// queries of one MFMA sweep: the tile-major kernel stages a slot's query
// limbs in LDS, which caps its columns; the direct kernel reads them from L2
// (256 columns cap the limb scratch at 128 queries per sweep either way)
int mfma_sweep_queries(const kspir_context* c, int n1) {
    const int lds_cap = 160 * 1024;
    int cmax = c->mfma_direct ? 256 : lds_cap / (4 * (n1 + 16)) / 16 * 16;
    if (cmax > 256) cmax = 256;
    return cmax / 2 < c->mfma_max_chunk ? cmax / 2 : c->mfma_max_chunk;
}

// k_db_mfma_direct in either first-dimension layout
template <int KS, int RTW>
void launch_direct(bool zmaj, dim3 g, hipStream_t st, const int8_t* dbl, const int8_t* ql, uint32_t* out, int R, int Rt,
                   int r0, int C, int n2, int nq, const QS& S, const MfmaConst& MC, const Tabs& T,
                   const DbSegs& SG = DbSegs{}) {
    if (zmaj)
        hipLaunchKernelGGL((k_db_mfma_direct<KS, RTW, true>), g, dim3(512), 0, st, dbl, ql, out, R, Rt, r0, C, n2, nq, S, MC, T, SG);
    else
        hipLaunchKernelGGL((k_db_mfma_direct<KS, RTW, false>), g, dim3(512), 0, st, dbl, ql, out, R, Rt, r0, C, n2, nq, S, MC, T, SG);
}

// Queries q_lo .. q_lo + nq (in sweeps of mfma_sweep_queries), database
// rows of blocks b0 .. b0 + nb (nb 0: every block); the products land in
// c->dbout, block-relative to b0.  A chunked pipeline calls it once with
// parts 1 (query limbs) and then per chunk with parts 6, always with the
// same nq and nb, so every call carves the scratch the same way.  zmaj
// (direct pass only): dbout in the slot-group-major layout of the chunked
// pipeline.
// the database's 7-bit int8 limbs, built once, by the first batched call
// of any context of its device
int ensure_limbs(kspir_database* db, hipStream_t st) {
    std::lock_guard<std::mutex> lock(db->limbs_mu);
    if (db->limbs) return KSPIR_OK;
    const int Rt = db->r * db->n2;
    int8_t* l = nullptr;
    HIP_OK(hipMalloc(&l, (size_t)db->r * kN * (kN / 2) * 8));
    hipLaunchKernelGGL(k_db_limbs, dim3(kN / 64, Rt), dim3(256), 0, st, db->words, l, db->n1, Rt);
    LAUNCH_OK("k_db_limbs");
    HIP_OK(hipStreamSynchronize(st));   // complete before another context's stream reads them
    db->limbs = l;
    return KSPIR_OK;
}

// the MFMA pass's scratch for nq queries of N1 n1 in sweeps of
// mfma_sweep_queries: the query limbs, and (tile-major pass) the products of
// R rows; grows on demand
int ensure_mws(kspir_context* c, int n1, int nq, int R, hipStream_t st, size_t* ql_bytes_out) {
    const int nqc_max = mfma_sweep_queries(c, n1);
    const int C_full = ((2 * (nq < nqc_max ? nq : nqc_max)) + 15) / 16 * 16;
    const size_t ql_bytes = (size_t)kN * 2 * 4 * C_full * n1;
    const size_t mf_bytes = c->mfma_direct ? 0 : (size_t)2 * R * C_full * kN * 4;
    if (ql_bytes + mf_bytes > c->mws_bytes) {
        if (c->mws) {
            HIP_OK(hipStreamSynchronize(st));
            HIP_OK(hipStreamSynchronize(c->stream));
            if (c->last_stream) HIP_OK(hipStreamSynchronize(c->last_stream));
            drop_graphs(c);
            HIP_OK(hipFree(c->mws));
            c->mws = nullptr;
        }
        HIP_OK(hipMalloc(&c->mws, ql_bytes + mf_bytes));
        c->mws_bytes = ql_bytes + mf_bytes;
    }
    *ql_bytes_out = ql_bytes;
    return KSPIR_OK;
}

// 2^(7w) * 2^32 mod q_p: the weights that recombine limb-pair sums into
// Montgomery-form residues
MfmaConst mfma_consts(const kspir_context* c) {
    MfmaConst MC;
    for (int p = 0; p < 2; ++p) {
        const u64 q = c->T.p[p].q;
        for (int w = 0; w < 7; ++w) MC.cw[p][w] = (uint32_t)(((u128)powmod(2, 7 * w, q) << 32) % q);
    }
    return MC;
}

// query limbs of nqc queries from their rotation words (C columns)
int launch_ct_limbs(int n1, const uint32_t* ctw, int8_t* ql, int nqc, int C, const QS& S, hipStream_t st) {
    switch (n1) {
        case 64: hipLaunchKernelGGL(k_ct_limbs<64>, dim3(kN / 32, C / 2, 2), dim3(256), 0, st, ctw, ql, nqc, C, S); break;
        case 128: hipLaunchKernelGGL(k_ct_limbs<128>, dim3(kN / 32, C / 2, 2), dim3(256), 0, st, ctw, ql, nqc, C, S); break;
        default: hipLaunchKernelGGL(k_ct_limbs<256>, dim3(kN / 32, C / 2, 2), dim3(256), 0, st, ctw, ql, nqc, C, S); break;
    }
    LAUNCH_OK("k_ct_limbs");
    return KSPIR_OK;
}

int launch_db_batched(kspir_context* c, const kspir_database* dbc, int q_lo, int nq, hipStream_t st, int parts = 7,
                      int b0 = 0, int nb = 0, bool zmaj = false) {
    kspir_database* db = const_cast<kspir_database*>(dbc);
    const int r = db->r, n1 = db->n1, n2 = db->n2, Rt = r * n2;
    if (nb <= 0) nb = r;
    const int R = nb * n2, r0 = b0 * n2;
    int rc = ensure_limbs(db, st);
    if (rc) return rc;
    const bool direct = c->mfma_direct != 0;
    const int nqc_max = mfma_sweep_queries(c, n1);
    size_t ql_bytes = 0;
    rc = ensure_mws(c, n1, nq, R, st, &ql_bytes);
    if (rc) return rc;
    int8_t* ql = (int8_t*)c->mws;
    uint32_t* mf = (uint32_t*)((char*)c->mws + ql_bytes);
    const MfmaConst MC = mfma_consts(c);
    const QS& S = c->S;
    for (int q0 = q_lo; q0 < q_lo + nq; q0 += nqc_max) {
        const int nqc = q_lo + nq - q0 < nqc_max ? q_lo + nq - q0 : nqc_max;
        const int C = (2 * nqc + 15) / 16 * 16;
        if (parts & 1) {
            rc = launch_ct_limbs(n1, c->ctw + q0 * S.ctw, ql, nqc, C, S, st);
            if (rc) return rc;
        }
        if (direct && (parts & 2)) {
            // products straight into dbout: 16 slots x (16 RTW) rows per workgroup
            const int rtw = (n1 <= 128 && R % 32 == 0) ? 2 : 1;
            const dim3 g((kN / 16) * (R / (16 * rtw)) * 2);
            uint32_t* out = c->dbout + q0 * S.dbout;
            if (n1 == 64 && rtw == 2) launch_direct<1, 2>(zmaj, g, st, db->limbs, ql, out, R, Rt, r0, C, n2, nqc, S, MC, c->T);
            else if (n1 == 64) launch_direct<1, 1>(zmaj, g, st, db->limbs, ql, out, R, Rt, r0, C, n2, nqc, S, MC, c->T);
            else if (n1 == 128 && rtw == 2) launch_direct<2, 2>(zmaj, g, st, db->limbs, ql, out, R, Rt, r0, C, n2, nqc, S, MC, c->T);
            else if (n1 == 128) launch_direct<2, 1>(zmaj, g, st, db->limbs, ql, out, R, Rt, r0, C, n2, nqc, S, MC, c->T);
            else launch_direct<4, 1>(zmaj, g, st, db->limbs, ql, out, R, Rt, r0, C, n2, nqc, S, MC, c->T);
            LAUNCH_OK("k_db_mfma_direct");
            continue;
        }
        if (direct) continue;
        const size_t lds = (size_t)4 * C * (n1 + 16);
        if (parts & 2) switch (n1) {
            case 64:
                HIP_OK(hipFuncSetAttribute((const void*)k_db_mfma<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                hipLaunchKernelGGL(k_db_mfma<1>, dim3(kN), dim3(256), lds, st, db->limbs, ql, mf, R, Rt, r0, C, MC, c->T);
                break;
            case 128:
                HIP_OK(hipFuncSetAttribute((const void*)k_db_mfma<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                hipLaunchKernelGGL(k_db_mfma<2>, dim3(kN), dim3(256), lds, st, db->limbs, ql, mf, R, Rt, r0, C, MC, c->T);
                break;
            default:
                HIP_OK(hipFuncSetAttribute((const void*)k_db_mfma<4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                hipLaunchKernelGGL(k_db_mfma<4>, dim3(kN), dim3(256), lds, st, db->limbs, ql, mf, R, Rt, r0, C, MC, c->T);
                break;
        }
        LAUNCH_OK("k_db_mfma");
        if (parts & 4)
            hipLaunchKernelGGL(k_mfma_scatter, dim3(kN / 32, (R / 16) * (C / 16), 2), dim3(256), 0, st, mf,
                               c->dbout + q0 * S.dbout, R, C, n2, nqc, S);
        LAUNCH_OK("k_mfma_scatter");
    }
    return KSPIR_OK;
}

// baby steps (twosteps.cpp:1206-1214, evalAutoSRNSSmallKeys) of nq queries
// of N1 n1: the rotation words ctw of every query
int enqueue_baby(kspir_context* c, const KeyPtrs* kp, int nq, int n1, const uint64_t* q_dev, hipStream_t st) {
    const int n1h = n1 / 2;
    const Tabs& T = c->T;
    const QS& S = c->S;
    hipLaunchKernelGGL(k_baby_prep, dim3(16, nq), dim3(kNttThreads), 0, st, q_dev, c->D0, c->bhat0, c->ctw, n1, S, T);
    LAUNCH_OK("k_baby_prep");
    // rotation N1/2-1 of the query first: its switched words and the
    // un-switched intermediate that the second half rotates
    hipLaunchKernelGGL(k_baby_rot, dim3(2 * nq), dim3(kNttThreads), 0, st, c->D0, c->bhat0, kp, c->ctw, c->mid, 1,
                       n1h - 2, nq, 1, n1h - 2, S.D, S.bhat, S, T);
    LAUNCH_OK("k_baby_rot(intermediate)");
    hipLaunchKernelGGL(k_baby_mid_coef, dim3(3, nq), dim3(kNttThreads), 0, st, c->mid, c->icoef, S, T);
    LAUNCH_OK("k_baby_mid_coef");
    hipLaunchKernelGGL(k_baby_mid_digits, dim3(9, nq), dim3(kNttThreads), 0, st, c->icoef, c->D1, S, T);
    LAUNCH_OK("k_baby_mid_digits");
    // the other rotations, both halves, one read of each key: rotations
    // 1 .. N1/2-2 of the query and N1/2 .. N1-1 (from the intermediate)
    hipLaunchKernelGGL(k_baby_rot_pair, dim3(n1h * 2 * nq), dim3(kNttThreads), 0, st, c->D0, c->bhat0, c->D1,
                       c->mid + 3 * kN, kp, c->ctw, n1h, nq, S, T);
    LAUNCH_OK("k_baby_rot_pair");
    return KSPIR_OK;
}

// packing tree on the NTT-domain planes at `cur` for nq queries; returns the
// buffer that holds the packed RLWE of each query (answer.cpp:800-835)
int enqueue_packing(kspir_context* c, const KeyPtrs* kp, int r, int nq, hipStream_t st, uint32_t** out) {
    const Tabs& T = c->T;
    const QS& S = c->S;
    uint32_t* cur = c->extA;
    uint32_t* nxt = c->extB;
    for (int i = r / 2; i > 0; i >>= 1) {
        const int li = ilog2(i);
        const uint32_t g = (uint32_t)(kN / i + 1);
        const uint32_t ginv = (uint32_t)powmod(g, kN - 1, 2 * kN);   // odd g has order | N in (Z/2N)*
        if (c->fused_pack && (c->cur_cb > 0 || nq >= c->fused_min_batch) && (int64_t)i * nq >= c->pack_fused_min_wgs) {
            hipLaunchKernelGGL(k_pack_level, dim3(i * nq), dim3(512), 0, st, cur, kp, li, nxt, i, g, ginv, S, T,
                               PackSegs{});
            LAUNCH_OK("k_pack_level");
        } else {
            hipLaunchKernelGGL(k_pack_coef, dim3(i, 2 * nq), dim3(kNttThreads), 0, st, cur, c->pcoef, i, S, T);
            LAUNCH_OK("k_pack_coef");
            hipLaunchKernelGGL(k_pack_digit, dim3(i, 4, 2 * nq), dim3(kNttThreads), 0, st, c->pcoef, kp, li,
                               c->ppart, ginv, S, T);
            LAUNCH_OK("k_pack_digit");
            hipLaunchKernelGGL(k_pack_combine, dim3(i, 2, 2 * nq), dim3(kNttThreads), 0, st, cur, c->ppart, nxt, i,
                               g, S, T);
            LAUNCH_OK("k_pack_combine");
        }
        uint32_t* t = cur; cur = nxt; nxt = t;
    }
    *out = cur;
    return KSPIR_OK;
}

// The batched pipeline after the baby steps, one chunk of c->cur_cb blocks at
// a time: per MFMA sweep of queries the query limbs once, then per chunk the
// database pass over the chunk's rows and the giant steps with the external
// products of every (query, block) of the chunk (k_giant_fused_dma<true, zmaj>,
// block offset b0 into ext); then the packing tree over all blocks.  The
// database is still read once per sweep, a chunk at a time.
int enqueue_chunked(kspir_context* c, const kspir_database* db, const KeyPtrs* kp, int nq, uint64_t* resp,
                    hipStream_t st) {
    const int r = db->r, n1 = db->n1, n2 = db->n2, cb = c->cur_cb;
    const QS& S = c->S;
    const Tabs& T = c->T;
    const int nqc_max = mfma_sweep_queries(c, n1);
    // the direct MFMA pass and the giant steps share the slot-group-major
    // first-dimension layout here (dbout is internal to this pipeline)
    const bool zmaj = c->mfma_direct != 0;
    for (int q0 = 0; q0 < nq; q0 += nqc_max) {
        const int nqc = nq - q0 < nqc_max ? nq - q0 : nqc_max;
        int rc = launch_db_batched(c, db, q0, nqc, st, 1, 0, cb);
        if (rc) return rc;
        for (int b0 = 0; b0 < r; b0 += cb) {
            rc = launch_db_batched(c, db, q0, nqc, st, 6, b0, cb, zmaj);
            if (rc) return rc;
            if (zmaj)
                hipLaunchKernelGGL((k_giant_fused_dma<true, true>), dim3(cb * nqc), dim3(kFusedThreads), 0, st,
                                   c->dbout + q0 * S.dbout, kp + q0, nullptr, c->extA + q0 * S.ext, cb, nqc, n1, n2, S,
                                   T, b0, GiantSegs{});
            else
                hipLaunchKernelGGL((k_giant_fused_dma<true, false>), dim3(cb * nqc), dim3(kFusedThreads), 0, st,
                                   c->dbout + q0 * S.dbout, kp + q0, nullptr, c->extA + q0 * S.ext, cb, nqc, n1, n2, S,
                                   T, b0, GiantSegs{});
            LAUNCH_OK("k_giant_fused_dma");
        }
    }
    uint32_t* packed = nullptr;
    int rc = enqueue_packing(c, kp, r, nq, st, &packed);
    if (rc) return rc;
    hipLaunchKernelGGL(k_response, dim3(2, nq), dim3(kNttThreads), 0, st, packed, resp, S.ext, T);
    LAUNCH_OK("k_response");
    return KSPIR_OK;
}

// A batch spanning several databases of one N1 (segment s: queries
// q_lo .. q_lo + nq of the batch, against database db), one launch per stage
// and chunk for all of them: the baby steps of every query; per MFMA sweep
// of queries their limbs; per chunk of cb blocks one database pass over the
// chunk's rows of every database that still has blocks there, then one
// giant-step launch over every (query, block) of that chunk; one packing
// launch per level over the queries of every database deep enough for it;
// one response launch.  Each segment's external products are written to the
// buffer that is the packing tree's input when its first level comes round
// (the levels ping-pong between extA and extB), so a database with smaller
// r joins the tree at its own depth with no copy.
struct Seg {
    const kspir_database* db;
    int q_lo, nq;
};

int enqueue_multi(kspir_context* c, const std::vector<Seg>& segs, const KeyPtrs* kp, int nq, const uint64_t* q_dev,
                  uint64_t* resp, hipStream_t st) {
    const int n1 = segs[0].db->n1, n2 = segs[0].db->n2, cb = c->cur_cb;
    int rmax = 0;
    for (const Seg& g : segs) rmax = g.db->r > rmax ? g.db->r : rmax;
    const int lmax = ilog2(rmax);
    const QS& S = c->S;
    const Tabs& T = c->T;
    int rc = enqueue_baby(c, kp, nq, n1, q_dev, st);
    if (rc) return rc;
    for (const Seg& g : segs) {
        rc = ensure_limbs(const_cast<kspir_database*>(g.db), st);
        if (rc) return rc;
    }
    size_t ql_bytes = 0;
    rc = ensure_mws(c, n1, nq, 0, st, &ql_bytes);
    if (rc) return rc;
    int8_t* ql = (int8_t*)c->mws;
    const MfmaConst MC = mfma_consts(c);
    const int nqc_max = mfma_sweep_queries(c, n1);
    for (int q0 = 0; q0 < nq; q0 += nqc_max) {
        const int nqc = nq - q0 < nqc_max ? nq - q0 : nqc_max;
        const int C = (2 * nqc + 15) / 16 * 16;
        rc = launch_ct_limbs(n1, c->ctw + q0 * S.ctw, ql, nqc, C, S, st);
        if (rc) return rc;
        for (int b0 = 0; b0 < rmax; b0 += cb) {
            // the segments with queries in this sweep and blocks in this chunk
            bool rtw2 = n1 <= 128;
            for (const Seg& g : segs) {
                const int lo = g.q_lo > q0 ? g.q_lo : q0, hi = g.q_lo + g.nq < q0 + nqc ? g.q_lo + g.nq : q0 + nqc;
                const int nb = g.db->r - b0 < cb ? g.db->r - b0 : cb;
                if (lo < hi && nb > 0 && (nb * n2) % 32) rtw2 = false;
            }
            const int rtw = rtw2 ? 2 : 1;
            DbSegs D{};
            GiantSegs G{};
            int dwg = 0, gwg = 0, k = 0;
            for (const Seg& g : segs) {
                const int lo = g.q_lo > q0 ? g.q_lo : q0, hi = g.q_lo + g.nq < q0 + nqc ? g.q_lo + g.nq : q0 + nqc;
                const int r = g.db->r, nb = r - b0 < cb ? r - b0 : cb;
                if (lo >= hi || nb <= 0) continue;
                const int R = nb * n2;
                D.dbl[k] = g.db->limbs;
                D.Rt[k] = r * n2;
                D.R[k] = R;
                D.r0[k] = b0 * n2;
                D.ct_lo[k] = 2 * (lo - q0) / 16;
                D.ct_hi[k] = (2 * (hi - q0) + 15) / 16;
                D.q_lo[k] = lo - q0;
                D.q_hi[k] = hi - q0;
                dwg += (kN / 16) * (R / (16 * rtw)) * 2;
                D.wg_end[k] = dwg;
                G.q_lo[k] = lo - q0;
                G.nb[k] = nb;
                G.out[k] = (((lmax - ilog2(r)) & 1) ? c->extB : c->extA) + q0 * S.ext;
                gwg += nb * (hi - lo);
                G.wg_end[k] = gwg;
                ++k;
            }
            if (!k) continue;
            D.n = G.n = k;
            uint32_t* out = c->dbout + q0 * S.dbout;
            const dim3 g(dwg);
            if (n1 == 64 && rtw == 2) launch_direct<1, 2>(true, g, st, nullptr, ql, out, 0, 0, 0, C, n2, nqc, S, MC, T, D);
            else if (n1 == 64) launch_direct<1, 1>(true, g, st, nullptr, ql, out, 0, 0, 0, C, n2, nqc, S, MC, T, D);
            else if (n1 == 128 && rtw == 2) launch_direct<2, 2>(true, g, st, nullptr, ql, out, 0, 0, 0, C, n2, nqc, S, MC, T, D);
            else if (n1 == 128) launch_direct<2, 1>(true, g, st, nullptr, ql, out, 0, 0, 0, C, n2, nqc, S, MC, T, D);
            else launch_direct<4, 1>(true, g, st, nullptr, ql, out, 0, 0, 0, C, n2, nqc, S, MC, T, D);
            LAUNCH_OK("k_db_mfma_direct(segments)");
            hipLaunchKernelGGL((k_giant_fused_dma<true, true>), dim3(gwg), dim3(kFusedThreads), 0, st, out, kp + q0,
                               nullptr, nullptr, 0, 0, n1, n2, S, T, b0, G);
            LAUNCH_OK("k_giant_fused_dma(segments)");
        }
    }
    // packing: level i runs for the queries of every database with r >= 2 i
    uint32_t* cur = c->extA;
    uint32_t* nxt = c->extB;
    for (int i = rmax / 2; i > 0; i >>= 1) {
        const int li = ilog2(i);
        const uint32_t g = (uint32_t)(kN / i + 1);
        const uint32_t ginv = (uint32_t)powmod(g, kN - 1, 2 * kN);
        PackSegs P{};
        int wg = 0, k = 0;
        for (const Seg& sg : segs) {
            if (sg.db->r < 2 * i) continue;
            P.q_lo[k] = sg.q_lo;
            wg += i * sg.nq;
            P.wg_end[k++] = wg;
        }
        P.n = k;
        hipLaunchKernelGGL(k_pack_level, dim3(wg), dim3(512), 0, st, cur, kp, li, nxt, i, g, ginv, S, T, P);
        LAUNCH_OK("k_pack_level(segments)");
        uint32_t* t = cur; cur = nxt; nxt = t;
    }
    hipLaunchKernelGGL(k_response, dim3(2, nq), dim3(kNttThreads), 0, st, cur, resp, S.ext, T);
    LAUNCH_OK("k_response");
    return KSPIR_OK;
}

// The online answer (test-pir.cpp:107-116) for nq queries against one
// database, enqueued on `st`.  q_dev: [nq][4][N] u64; resp: [nq][2][N] u64
// (unused when first_dim_only).  kp: device array of nq key-pointer sets.
// Optional inspection outputs for query 0: fd_u64 [r][2][N] (NTT form),
// ext_u64 [r][2][N] (coefficient form).
int enqueue_answer(kspir_context* c, const kspir_database* db, const KeyPtrs* kp, int nq, const uint64_t* q_dev,
                   uint64_t* resp, hipStream_t st, uint64_t* fd_u64, uint64_t* ext_u64, bool first_dim_only,
                   bool allow_mfma = true) {
    const int r = db->r, n1 = db->n1, n2 = db->n2;
    const Tabs& T = c->T;
    const QS& S = c->S;
    int rc0 = enqueue_baby(c, kp, nq, n1, q_dev, st);
    if (rc0) return rc0;
    // database pass (crt.cpp:643-768): int8-MFMA limb GEMM for a batch,
    // u64-accumulating VALU stream for a single query
    if (c->cur_cb > 0) {
        if (!allow_mfma || first_dim_only)
            return fail(KSPIR_EINVAL, "internal: chunked workspace for an unchunked pipeline");
        return enqueue_chunked(c, db, kp, nq, resp, st);
    }
    // a database kept as limbs only (kspir_database_compact) always takes the MFMA pass
    const bool mfma = allow_mfma && mfma_shape(db) && (nq >= c->mfma_min_batch || !db->words);
    if (!mfma && !db->words)
        return fail(KSPIR_EINVAL, "database kept as int8 limbs only (kspir_database_compact): this call needs its words");
    int rc = mfma ? launch_db_batched(c, db, 0, nq, st) : launch_db_product(c, db, nq, st);
    if (rc) return rc;
    // giant steps (twosteps.cpp:599-677) and the first-dimension sum: one
    // fused workgroup per (query, block) for a batch; for few queries, one
    // workgroup per (block, row, prime) keeps the GPU full
    bool ext_done = false;   // the giant-step kernel ran the external products too
    if (nq >= c->fused_min_batch) {
        if (c->giant_dma && c->giant_wgs_per_cu == 1 && c->ext_fused && !first_dim_only) {
            // the ext output of a query has the ecoef stride's layout [r][ab][p][N]
            hipLaunchKernelGGL((k_giant_fused_dma<true, false>), dim3(r * nq), dim3(kFusedThreads), 0, st, c->dbout, kp,
                               fd_u64 ? c->fd : nullptr, c->extA, r, nq, n1, n2, S, T, 0, GiantSegs{});
            ext_done = true;
        } else if (c->giant_dma && c->giant_wgs_per_cu == 1)
            hipLaunchKernelGGL((k_giant_fused_dma<false, false>), dim3(r * nq), dim3(kFusedThreads), 0, st, c->dbout, kp,
                               fd_u64 ? c->fd : nullptr, c->ecoef, r, nq, n1, n2, S, T, 0, GiantSegs{});
        else if (c->giant_wgs_per_cu == 2)
            hipLaunchKernelGGL(k_giant_fused2, dim3(r * nq), dim3(kFusedThreads), 0, st, c->dbout, kp, c->fd, c->ecoef,
                               r, nq, n1, n2, S, T);
        else
            hipLaunchKernelGGL(k_giant_fused, dim3(r * nq), dim3(kFusedThreads), 0, st, c->dbout, kp, c->fd, c->ecoef,
                               r, nq, n1, n2, S, T);
        LAUNCH_OK("k_giant_fused");
    } else {
        if (n2 > 1) {
            hipLaunchKernelGGL(k_giant_coef, dim3(n2 - 1, r, 2 * nq), dim3(kNttThreads), 0, st, c->dbout, c->gcoef,
                               n2, S, T);
            LAUNCH_OK("k_giant_coef");
            hipLaunchKernelGGL(k_giant_ks, dim3(n2 - 1, r, 2 * nq), dim3(kNttThreads), 0, st, c->dbout, c->gcoef, kp,
                               c->gks, n1, n2, S, T);
            LAUNCH_OK("k_giant_ks");
        }
        // the sum over giant rows on every lane of the GPU, then the INTTs
        hipLaunchKernelGGL(k_fd_sum, dim3(grid1d((size_t)nq * r * 4 * kN)), dim3(256), 0, st, c->dbout, c->gks, c->fd,
                           r, n2, nq, S, T);
        LAUNCH_OK("k_fd_sum");
        hipLaunchKernelGGL(k_fd_intt, dim3(r, 2, 2 * nq), dim3(kNttThreads), 0, st, c->fd, c->ecoef, S, T);
        LAUNCH_OK("k_fd_intt");
    }
    if (fd_u64) {
        hipLaunchKernelGGL(k_rns_to_u64, dim3(grid1d((size_t)r * 2 * kN)), dim3(256), 0, st, c->fd, fd_u64, r, T);
        LAUNCH_OK("k_rns_to_u64");
    }
    if (first_dim_only) return KSPIR_OK;
    // external products (lwe.cpp:761-808)
    if (!ext_done) {
        hipLaunchKernelGGL(k_ext_digit, dim3(r, 4, 2 * nq), dim3(kNttThreads), 0, st, c->ecoef, kp, c->epart, S, T);
        LAUNCH_OK("k_ext_digit");
        hipLaunchKernelGGL(k_ext_sum, dim3(grid1d((size_t)nq * r * 4 * kN)), dim3(256), 0, st, c->epart, c->extA, r,
                           nq, S, T);
        LAUNCH_OK("k_ext_sum");
    }
    if (ext_u64) {
        hipLaunchKernelGGL(k_response, dim3(2, r), dim3(kNttThreads), 0, st, c->extA, ext_u64, (size_t)4 * kN, T);
        LAUNCH_OK("k_response(ext)");
    }
    // packing (answer.cpp:800-835)
    uint32_t* packed = nullptr;
    rc = enqueue_packing(c, kp, r, nq, st, &packed);
    if (rc) return rc;
    hipLaunchKernelGGL(k_response, dim3(2, nq), dim3(kNttThreads), 0, st, packed, resp, S.ext, T);
    LAUNCH_OK("k_response");
    return KSPIR_OK;
}

// the database and keys fit together; the keys hold what the call needs
int check_pair(const kspir_context* c, const kspir_database* db, const kspir_keys* k, bool first_dim_only = false) {
    if (!c || !db || !k) return fail(KSPIR_EINVAL, "null handle");
    if (db->ctx->device != c->device || k->ctx->device != c->device)
        return fail(KSPIR_EINVAL, "handle belongs to another device");
    if (db->r != k->r || db->n1 != k->n1)
        return fail(KSPIR_EINVAL, "database and keys disagree on (r, N1)");
    if (!k->baby || (db->n2 > 1 && !k->giant)) return fail(KSPIR_EINVAL, "keys lack baby- or giant-step keys");
    if (!first_dim_only && (!k->rgsw || (k->r > 1 && !k->packing)))
        return fail(KSPIR_EINVAL, "keys lack the RGSW ciphertext or packing keys");
    return KSPIR_OK;
}

// enqueue the answer pipeline through a cached hipGraph: the first call with
// a given (database, key table, batch size, buffers, stream) runs eagerly
// (which also makes every lazy allocation) and then captures the same
// launches; later calls replay the graph
template <class Enq>
int run_graph(kspir_context* c, const kspir_database* db, const std::vector<const void*>& segs, const KeyPtrs* kp,
              int nq, const uint64_t* q, uint64_t* resp, hipStream_t st, Enq enq) {
    if (st != c->stream) c->last_stream = st;
    if (!c->use_graphs) return enq();
    bool found = false;
    hipGraphExec_t exec = nullptr;
    {
        std::lock_guard<std::mutex> lock(c->graph_mu);
        for (auto& g : c->graphs)
            if (g.db == db && g.kp == kp && g.nq == nq && g.q == q && g.resp == resp && g.st == st && g.segs == segs) {
                found = true;
                exec = g.exec;
                break;
            }
    }
    if (found) {
        if (!exec) return enq();
        HIP_OK(hipGraphLaunch(exec, st));
        return KSPIR_OK;
    }
    int rc = enq();
    if (rc) return rc;
    kspir_context::Graph g{db, kp, nq, q, resp, st, nullptr, segs};
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) == hipSuccess) {
        const int crc = enq();
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (crc == KSPIR_OK && e == hipSuccess && graph) {
            if (hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) != hipSuccess) g.exec = nullptr;
        }
        if (graph) hipGraphDestroy(graph);
    }
    (void)hipGetLastError();   // a failed capture leaves the eager path in place
    if (c->graphs.size() >= 64) {
        HIP_OK(hipStreamSynchronize(st));   // earlier replays may still be running
        drop_graphs(c);
    }
    std::lock_guard<std::mutex> lock(c->graph_mu);
    c->graphs.push_back(g);
    return KSPIR_OK;
}

int run_answer(kspir_context* c, const kspir_database* db, const KeyPtrs* kp, int nq, const uint64_t* q,
               uint64_t* resp, hipStream_t st) {
    return run_graph(c, db, {}, kp, nq, q, resp, st,
                     [&] { return enqueue_answer(c, db, kp, nq, q, resp, st, nullptr, nullptr, false); });
}

// device key-pointer table for a batch; re-uploaded only when it changes
int stage_batch_keys(kspir_context* c, const kspir_keys* const* keys, int nq) {
    if (nq > c->kp_cap) {
        HIP_OK(hipStreamSynchronize(c->stream));
        if (c->last_stream) HIP_OK(hipStreamSynchronize(c->last_stream));
        drop_graphs(c);
        if (c->kp_dev) HIP_OK(hipFree(c->kp_dev));
        if (c->kp_host) HIP_OK(hipHostFree(c->kp_host));
        c->kp_dev = nullptr;
        c->kp_host = nullptr;
        HIP_OK(hipMalloc(&c->kp_dev, sizeof(KeyPtrs) * nq));
        HIP_OK(hipHostMalloc(&c->kp_host, sizeof(KeyPtrs) * nq, hipHostMallocDefault));
        c->kp_cap = nq;
        c->kp_count = 0;
    }
    bool same = c->kp_count == nq;
    for (int i = 0; i < nq && same; ++i)
        same = std::memcmp(&c->kp_host[i], &keys[i]->ptrs, sizeof(KeyPtrs)) == 0;
    if (same) return KSPIR_OK;
    HIP_OK(hipStreamSynchronize(c->stream));   // the previous table may still be in flight
    if (c->last_stream) HIP_OK(hipStreamSynchronize(c->last_stream));
    for (int i = 0; i < nq; ++i) c->kp_host[i] = keys[i]->ptrs;
    HIP_OK(hipMemcpyAsync(c->kp_dev, c->kp_host, sizeof(KeyPtrs) * nq, hipMemcpyHostToDevice, c->stream));
    c->kp_count = nq;
    return KSPIR_OK;
}

// blocks per chunk of a batch spanning databases: the context's chunk size,
// widened to whole 16-row MFMA tiles, at most the largest database's r
int multi_chunk_blocks(const kspir_context* c, int rmax, int n2) {
    int cb = c->chunk_blocks > 0 && c->chunk_blocks < rmax ? c->chunk_blocks : rmax;
    while ((cb * n2) % 16) cb *= 2;
    return cb > rmax ? rmax : cb;
}

// Device bytes of a batch of nq queries spanning databases of N1 n1 whose
// largest has rmax blocks: the workspace carved for that shape at the total
// count, in chunks, and the query limbs of one sweep
size_t multi_footprint_bytes(kspir_context* c, int rmax, int n1, int nq) {
    const int n2 = kN / 2 / n1;
    const int nqc = nq < mfma_sweep_queries(c, n1) ? nq : mfma_sweep_queries(c, n1);
    return layout_workspace(c, rmax, n1, nq, nullptr, false, multi_chunk_blocks(c, rmax, n2)) * 4 +
           (size_t)kN * 2 * 4 * ((2 * nqc + 15) / 16 * 16) * n1;
}

// whether the context's options run batches spanning databases
bool multi_options(const kspir_context* c) {
    return c->mfma_direct && c->giant_dma && c->giant_wgs_per_cu == 1 && c->ext_fused;
}

// checks a batch spanning databases (segment s: seg_nq[s] consecutive
// queries against dbs[s], keys[q] per query), fills segs, carves the
// workspace (per-query buffers sized for the deepest database) and stages
// the key table
int prepare_multi(kspir_context* c, int nseg, const kspir_database* const* dbs, const int32_t* seg_nq,
                  const kspir_keys* const* keys, std::vector<Seg>& segs, std::vector<const void*>& sig, int* nq_out) {
    if (!c || !dbs || !seg_nq || !keys) return fail(KSPIR_EINVAL, "null argument");
    if (nseg < 1 || nseg > kMaxSeg)
        return fail(KSPIR_EINVAL, "a batch spans 1 to " + std::to_string(kMaxSeg) + " databases");
    if (!multi_options(c))
        return fail(KSPIR_EINVAL, "a batch spanning databases runs the chunked int8-MFMA pipeline: options mfma_direct, "
                                  "giant_dma, ext_fused on and giant_wgs_per_cu 1");
    segs.clear();
    sig.clear();
    int nq = 0, rmax = 0;
    for (int s = 0; s < nseg; ++s) {
        const kspir_database* db = dbs[s];
        if (!db) return fail(KSPIR_EINVAL, "null database handle");
        if (db->n1 != dbs[0]->n1) return fail(KSPIR_EINVAL, "the databases of one batch share N1");
        if (!mfma_shape(db)) return fail(KSPIR_EINVAL, "database shape has no int8-MFMA pass (N1 64..256, r*N2 % 16 == 0)");
        if (seg_nq[s] < 1) return fail(KSPIR_EINVAL, "every database of the batch needs at least one query");
        for (int q = nq; q < nq + seg_nq[s]; ++q) {
            int rc = check_pair(c, db, keys[q]);
            if (rc) return rc;
        }
        segs.push_back(Seg{db, nq, seg_nq[s]});
        sig.push_back(db);
        sig.push_back((const void*)(uintptr_t)seg_nq[s]);
        nq += seg_nq[s];
        rmax = db->r > rmax ? db->r : rmax;
    }
    const int n1 = dbs[0]->n1, n2 = dbs[0]->n2;
    const int cb = multi_chunk_blocks(c, rmax, n2);
    HIP_OK(hipSetDevice(c->device));
    int rc = carve_workspace_cb(c, rmax, n1, nq, cb);
    if (rc) return rc;
    rc = stage_batch_keys(c, keys, nq);
    if (rc) return rc;
    *nq_out = nq;
    return KSPIR_OK;
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

const char* kspir_last_error(void) { return g_err.c_str(); }

// error text for the library's other translation units (not exported)
__attribute__((visibility("hidden"))) int kspir_set_error(int code, const char* msg) { return fail(code, msg); }

int kspir_compute_permutation_matrix(int32_t* out, int32_t n1) {
    if (!out || n1 < 1 || n1 > kN / 2) return fail(KSPIR_EINVAL, "the row count must be in [1, N/2]");
    // slot s holds the evaluation at psi^e with e = 2*bitrev(s)+1; tau_g maps
    // it to the slot holding psi^(e*g mod 2N)
    std::vector<int32_t> slot_of(kN);
    for (int s = 0; s < kN; ++s) slot_of[(2 * bitrev12(s) + 1) >> 1] = s;
    uint32_t g = 1;
    for (int i = 0; i < n1; ++i) {
        for (int s = 0; s < kN; ++s) {
            const uint32_t e = ((2 * bitrev12(s) + 1) * g) & (2 * kN - 1);
            out[(size_t)i * kN + s] = slot_of[e >> 1];
        }
        g = (g * 5u) & (2 * kN - 1);
    }
    return KSPIR_OK;
}

int kspir_device_count(int* count) {
    HIP_OK(hipGetDeviceCount(count));
    return KSPIR_OK;
}

int kspir_context_create(int device, kspir_context** out) {
    if (!out) return fail(KSPIR_EINVAL, "null out");
    int n = 0;
    HIP_OK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(KSPIR_EINVAL, "no such device");
    if (device >= kMaxDevices) return fail(KSPIR_EINVAL, "device index above the library's limit of 64");
    HIP_OK(hipSetDevice(device));
    {
        // this device's copy of the no-allocation scratch array
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[device]) {
            void* p = nullptr;
            HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            if (!p) return fail(KSPIR_EHIP, "hipGetSymbolAddress(scratch) returned a null address");
            g_scratch_ptr[device] = (uint32_t*)p;
        }
    }
    kspir_context* c = new kspir_context();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail_hip("hipStreamCreate", e); }
    {
        std::lock_guard<std::mutex> lock(g_reg_mu);
        g_reg.push_back(c);
    }
    int rc = build_tables(c);
    if (rc != KSPIR_OK) { kspir_context_destroy(c); return rc; }
    *out = c;
    return KSPIR_OK;
}

void kspir_context_destroy(kspir_context* c) {
    if (!c) return;
    {
        std::lock_guard<std::mutex> lock(g_reg_mu);
        for (size_t i = 0; i < g_reg.size(); ++i)
            if (g_reg[i] == c) { g_reg.erase(g_reg.begin() + i); break; }
    }
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    drop_graphs(c);
    if (c->rng_jump) hipFree(c->rng_jump);
    if (c->enc_tab) hipFree(c->enc_tab);
    for (void* p : c->tables) hipFree(p);
    if (c->ws) hipFree(c->ws);
    if (c->kp_dev) hipFree(c->kp_dev);
    if (c->kp_host) hipHostFree(c->kp_host);
    if (c->mws) hipFree(c->mws);
    if (c->ows) hipFree(c->ows);
    if (c->ohost) hipHostFree(c->ohost);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

void* kspir_context_stream(kspir_context* c) { return c ? (void*)c->stream : nullptr; }

int kspir_database_create(kspir_context* c, const uint64_t* datacrt, int32_t r, int32_t n1, int on_device,
                          kspir_database** out) {
    if (!c || !datacrt || !out) return fail(KSPIR_EINVAL, "null argument");
    if (!pow2(r) || r > kN || !pow2(n1) || n1 < 32 || n1 > kN / 2)
        return fail(KSPIR_EINVAL, "r must be a power of two <= N, N1 a power of two in [32, N/2]");
    HIP_OK(hipSetDevice(c->device));
    const int n2 = kN / 2 / n1;
    const size_t words = (size_t)kN * kN / 2;
    uint64_t* dst = nullptr;
    HIP_OK(hipMalloc(&dst, words * 8 * (size_t)r));
    uint64_t* stage = nullptr;
    if (!on_device) {
        hipError_t e = hipMalloc(&stage, words * 8);
        if (e != hipSuccess) { hipFree(dst); return fail_hip("hipMalloc(stage)", e); }
    }
    for (int b = 0; b < r; ++b) {
        const uint64_t* src = datacrt + words * b;
        if (!on_device) {
            hipError_t e = hipMemcpyAsync(stage, src, words * 8, hipMemcpyHostToDevice, c->stream);
            if (e != hipSuccess) { hipFree(dst); hipFree(stage); return fail_hip("hipMemcpy(db)", e); }
            src = stage;
        }
        hipLaunchKernelGGL(k_db_transpose, dim3(kN / 32, n2 * n1 / 32), dim3(256), 0, c->stream, dst + words * b,
                           src, n2 * n1);
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) { hipFree(dst); if (stage) hipFree(stage); return fail_hip("k_db_transpose", e); }
    }
    if (stage) hipFree(stage);
    *out = new kspir_database{c, r, n1, n2, dst};
    return KSPIR_OK;
}

// gather table of database_tobsgsntt for N1 rotations: gtab[i][x] =
// dec[perm_i[x]], where perm_i is row i of compute_permutation_matrix and dec
// inverts compute_query_encode (utils.cpp:1657-1675): enc[i] =
// bitrev(rot_i >> 1) with rot_i = 5^i mod 2N for i < N/2 and
// rot_{N-1-i} = (2N-1) * 5^i mod 2N
static int db_gather_table(int n1, std::vector<uint16_t>& gtab) {
    std::vector<int32_t> perm((size_t)n1 * kN);
    int rc = kspir_compute_permutation_matrix(perm.data(), n1);
    if (rc) return rc;
    std::vector<uint32_t> rot(kN), dec(kN);
    rot[0] = 1;
    rot[kN - 1] = 2 * kN - 1;
    for (int i = 1; i < kN / 2; ++i) {
        rot[i] = rot[i - 1] * 5u % (2 * kN);
        rot[kN - 1 - i] = rot[kN - i] * 5u % (2 * kN);
    }
    for (int i = 0; i < kN; ++i) dec[bitrev12(rot[i] >> 1)] = (uint32_t)i;
    gtab.resize((size_t)n1 * kN);
    for (size_t t = 0; t < gtab.size(); ++t) gtab[t] = (uint16_t)dec[perm[t]];
    return KSPIR_OK;
}

int kspir_database_create_from_records(kspir_context* c, const uint64_t* records, int64_t block_stride, int32_t r,
                                       int32_t n1, int on_device, kspir_database** out) {
    if (!c || !records || !out) return fail(KSPIR_EINVAL, "null argument");
    if (!pow2(r) || r > kN || !pow2(n1) || n1 < 32 || n1 > kN / 2)
        return fail(KSPIR_EINVAL, "r must be a power of two <= N, N1 a power of two in [32, N/2]");
    if (block_stride < 0) return fail(KSPIR_EINVAL, "negative block stride");
    HIP_OK(hipSetDevice(c->device));
    const int n2 = kN / 2 / n1;
    const size_t words = (size_t)kN * kN / 2;
    std::vector<uint16_t> gh;
    int rc = db_gather_table(n1, gh);
    if (rc) return rc;
    uint64_t* dst = nullptr;
    uint64_t* stage = nullptr;
    uint16_t* gtab = nullptr;
    auto cleanup = [&]() { if (dst) hipFree(dst); if (stage) hipFree(stage); if (gtab) hipFree(gtab); };
    hipError_t e = hipMalloc(&dst, words * 8 * (size_t)r);
    if (e == hipSuccess) e = hipMalloc(&gtab, gh.size() * 2);
    if (e == hipSuccess && !on_device) e = hipMalloc(&stage, words * 8);
    if (e == hipSuccess) e = hipMemcpyAsync(gtab, gh.data(), gh.size() * 2, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) { cleanup(); return fail_hip("database_create_from_records", e); }
    for (int b = 0; b < r; ++b) {
        const uint64_t* src = records + (size_t)block_stride * b;
        if (!on_device) {
            if (b == 0 || block_stride != 0) {
                e = hipMemcpyAsync(stage, src, words * 8, hipMemcpyHostToDevice, c->stream);
                if (e != hipSuccess) { cleanup(); return fail_hip("hipMemcpy(records)", e); }
            }
            src = stage;
        }
        hipLaunchKernelGGL(k_db_prep, dim3(n2 * n1), dim3(kNttThreads), 0, c->stream, src, gtab, dst + words * b, n1,
                           1, c->T);
        e = hipGetLastError();
        if (e == hipSuccess && !on_device) e = hipStreamSynchronize(c->stream);   // stage reused next block
        if (e != hipSuccess) { cleanup(); return fail_hip("k_db_prep", e); }
    }
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { cleanup(); return fail_hip("k_db_prep", e); }
    if (stage) hipFree(stage);
    hipFree(gtab);
    *out = new kspir_database{c, r, n1, n2, dst};
    return KSPIR_OK;
}

int kspir_inverse_encode(kspir_context* c, uint64_t* messages, int32_t count) {
    if (!c || !messages || count < 0) return fail(KSPIR_EINVAL, "null argument or negative count");
    if (count == 0) return KSPIR_OK;
    const size_t words = (size_t)count * kN;
    for (size_t i = 0; i < words; ++i)
        if (messages[i] >= kP) return fail(KSPIR_EINVAL, "inverse_encode: record values must be below p = 65537");
    HIP_OK(hipSetDevice(c->device));
    // compute_query_decode (utils.cpp:1677-1685): the inverse of the slot
    // order db_gather_table uses, dec[bitrev(rot_i >> 1)] = i
    std::vector<uint16_t> dec(kN);
    {
        std::vector<uint32_t> rot(kN);
        rot[0] = 1;
        rot[kN - 1] = 2 * kN - 1;
        for (int i = 1; i < kN / 2; ++i) {
            rot[i] = rot[i - 1] * 5u % (2 * kN);
            rot[kN - 1 - i] = rot[kN - i] * 5u % (2 * kN);
        }
        for (int i = 0; i < kN; ++i) dec[bitrev12(rot[i] >> 1)] = (uint16_t)i;
    }
    uint64_t* d = nullptr;
    uint16_t* dd = nullptr;
    hipError_t e = hipMalloc(&d, words * 8);
    if (e == hipSuccess) e = hipMalloc(&dd, kN * 2);
    if (e == hipSuccess) e = hipMemcpyAsync(dd, dec.data(), kN * 2, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d, messages, words * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_inverse_encode, dim3(count), dim3(kNttThreads), 0, c->stream, d, dd, c->T);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(messages, d, words * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (d) hipFree(d);
    if (dd) hipFree(dd);
    if (e != hipSuccess) return fail_hip("inverse_encode", e);
    return KSPIR_OK;
}

int kspir_database_tobsgsntt(kspir_context* c, const uint64_t* data, int32_t n1, uint64_t* result) {
    if (!c || !data || !result) return fail(KSPIR_EINVAL, "null argument");
    if (!pow2(n1) || n1 < 1 || n1 > kN / 2) return fail(KSPIR_EINVAL, "N1 must be a power of two <= N/2");
    HIP_OK(hipSetDevice(c->device));
    const int n2 = kN / 2 / n1;
    const size_t words = (size_t)kN * kN / 2;
    std::vector<uint16_t> gh;
    int rc = db_gather_table(n1, gh);
    if (rc) return rc;
    uint64_t *src = nullptr, *dst = nullptr;
    uint16_t* gtab = nullptr;
    hipError_t e = hipMalloc(&src, words * 8);
    if (e == hipSuccess) e = hipMalloc(&dst, words * 8);
    if (e == hipSuccess) e = hipMalloc(&gtab, gh.size() * 2);
    if (e == hipSuccess) e = hipMemcpyAsync(gtab, gh.data(), gh.size() * 2, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(src, data, words * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_db_prep, dim3(n2 * n1), dim3(kNttThreads), 0, c->stream, src, gtab, dst, n1, 0, c->T);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(result, dst, words * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (src) hipFree(src);
    if (dst) hipFree(dst);
    if (gtab) hipFree(gtab);
    if (e != hipSuccess) return fail_hip("database_tobsgsntt", e);
    return KSPIR_OK;
}

int kspir_database_export(const kspir_database* db, uint64_t* datacrt) {
    if (!db || !datacrt) return fail(KSPIR_EINVAL, "null argument");
    kspir_context* c = db->ctx;
    HIP_OK(hipSetDevice(c->device));
    const size_t words = (size_t)kN * kN / 2;
    uint64_t* stage = nullptr;
    uint64_t* wstage = nullptr;   // a block's words rebuilt from the limbs (compacted database)
    HIP_OK(hipMalloc(&stage, words * 8));
    if (!db->words) {
        hipError_t e = hipMalloc(&wstage, words * 8);
        if (e != hipSuccess) { hipFree(stage); return fail_hip("hipMalloc(export)", e); }
    }
    hipError_t e = hipSuccess;
    for (int b = 0; b < db->r && e == hipSuccess; ++b) {
        const uint64_t* src = db->words ? db->words + words * b : wstage;
        if (!db->words) {
            hipLaunchKernelGGL(k_db_unlimbs, dim3(kN / 64, db->n2), dim3(256), 0, c->stream, db->limbs, wstage, db->n1,
                               db->r * db->n2, b * db->n2);
            e = hipGetLastError();
            if (e != hipSuccess) break;
        }
        hipLaunchKernelGGL(k_db_untranspose, dim3(kN / 32, db->n2 * db->n1 / 32), dim3(256), 0, c->stream, stage,
                           src, db->n2 * db->n1);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(datacrt + words * b, stage, words * 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    hipFree(stage);
    if (wstage) hipFree(wstage);
    if (e != hipSuccess) return fail_hip("database_export", e);
    return KSPIR_OK;
}

// Persisted database file: a 64-byte header, then the r blocks of words in
// the reference datacrt layout (what kspir_database_export returns).
//   0  char[8]  "KSPIRDB\0"
//   8  u32      format version (1)
//  12  u32      N
//  16  u32      r
//  20  u32      N1
//  24  u64      words per block (N * N / 2)
//  32  u64      FNV-1a 64 of the words (little-endian bytes)
//  40  u8[24]   zero
namespace {
constexpr char kDbMagic[8] = {'K', 'S', 'P', 'I', 'R', 'D', 'B', '\0'};
constexpr uint32_t kDbVersion = 1;
struct DbFileHeader {
    char magic[8];
    uint32_t version, n, r, n1;
    uint64_t words_per_block, fnv;
    uint8_t zero[24];
};
static_assert(sizeof(DbFileHeader) == 64, "database file header is 64 bytes");
static uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
    const uint8_t* b = (const uint8_t*)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ULL;
    return h;
}
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ULL;
}  // namespace

int kspir_database_save(const kspir_database* db, const char* path) {
    if (!db || !path) return fail(KSPIR_EINVAL, "null argument");
    const size_t words = (size_t)kN * kN / 2;
    std::vector<uint64_t> host(words * db->r);
    int rc = kspir_database_export(db, host.data());
    if (rc) return rc;
    DbFileHeader h{};
    std::memcpy(h.magic, kDbMagic, 8);
    h.version = kDbVersion;
    h.n = kN;
    h.r = (uint32_t)db->r;
    h.n1 = (uint32_t)db->n1;
    h.words_per_block = words;
    h.fnv = fnv1a(kFnvBasis, host.data(), host.size() * 8);
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(KSPIR_EINVAL, std::string("cannot open ") + path + " for writing");
    const bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(host.data(), 8, host.size(), f) == host.size();
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) return fail(KSPIR_EINVAL, std::string("short write to ") + path);
    return KSPIR_OK;
}

int kspir_database_load(kspir_context* c, const char* path, kspir_database** out) {
    if (!c || !path || !out) return fail(KSPIR_EINVAL, "null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(KSPIR_EINVAL, std::string("cannot open ") + path);
    DbFileHeader h{};
    if (std::fread(&h, sizeof h, 1, f) != 1) {
        std::fclose(f);
        return fail(KSPIR_EINVAL, std::string(path) + ": no database header");
    }
    const size_t words = (size_t)kN * kN / 2;
    if (std::memcmp(h.magic, kDbMagic, 8) != 0 || h.version != kDbVersion) {
        std::fclose(f);
        return fail(KSPIR_EINVAL, std::string(path) + ": not a database file of this format version");
    }
    if (h.n != (uint32_t)kN || h.words_per_block != words || h.r == 0 || h.r > (uint32_t)kN || h.n1 == 0) {
        std::fclose(f);
        return fail(KSPIR_EINVAL, std::string(path) + ": header shape (N, r, N1, block words) does not fit N = 4096");
    }
    std::vector<uint64_t> host(words * h.r);
    const size_t got = std::fread(host.data(), 8, host.size(), f);
    const bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    if (got != host.size() || extra)
        return fail(KSPIR_EINVAL, std::string(path) + ": file length does not match its header's r");
    if (fnv1a(kFnvBasis, host.data(), host.size() * 8) != h.fnv)
        return fail(KSPIR_EINVAL, std::string(path) + ": checksum mismatch (file corrupted)");
    return kspir_database_create(c, host.data(), (int32_t)h.r, (int32_t)h.n1, 0, out);
}

int kspir_database_compact(kspir_database* db) {
    if (!db) return fail(KSPIR_EINVAL, "null database handle");
    if (!db->words) return KSPIR_OK;
    if (!mfma_shape(db))
        return fail(KSPIR_EINVAL, "this database shape has no int8-MFMA pass (N1 64..256, r*N2 % 16 == 0): its words stay");
    kspir_context* c = db->ctx;
    HIP_OK(hipSetDevice(c->device));
    int rc = ensure_limbs(db, c->stream);
    if (rc) return rc;
    // graphs captured with the word pass name the words
    drop_graphs_everywhere(c->device, db);
    HIP_OK(hipFree(db->words));
    db->words = nullptr;
    return KSPIR_OK;
}

void kspir_database_destroy(kspir_database* db) {
    if (!db) return;
    hipSetDevice(db->ctx->device);
    drop_graphs_everywhere(db->ctx->device, db);
    hipFree(db->words);
    if (db->limbs) hipFree(db->limbs);
    delete db;
}

int kspir_keys_create(kspir_context* c, int32_t r, int32_t n1, const uint64_t* baby, const uint64_t* giant,
                      const uint64_t* packing, const uint64_t* rgsw, const int32_t* permutations, kspir_keys** out) {
    if (!c || !out) return fail(KSPIR_EINVAL, "null argument");
    if (!pow2(r) || r > kN || !pow2(n1) || n1 < 32 || n1 > kN / 2)
        return fail(KSPIR_EINVAL, "r must be a power of two <= N, N1 a power of two in [32, N/2]");
    const int n2 = kN / 2 / n1;
    if (!baby != !permutations) return fail(KSPIR_EINVAL, "baby-step keys and permutations go together");
    if (baby && n2 > 1 && !giant) return fail(KSPIR_EINVAL, "giant-step keys required when N2 > 1");
    std::vector<uint16_t> p16(permutations ? (size_t)n1 * kN : 0);
    for (size_t i = 0; i < p16.size(); ++i) {
        if (permutations[i] < 0 || permutations[i] >= kN) return fail(KSPIR_EINVAL, "permutation entry out of range");
        p16[i] = (uint16_t)permutations[i];
    }
    // the baby-step keys are stored through rows 1 .. N1/2, which must
    // therefore be permutations (compute_permutation_matrix's rows are)
    for (int row = 1; row <= n1 / 2 && !p16.empty(); ++row) {
        std::vector<uint8_t> seen(kN, 0);
        for (int x = 0; x < kN; ++x) {
            uint8_t& f = seen[p16[(size_t)row * kN + x]];
            if (f) return fail(KSPIR_EINVAL, "permutation row is not a permutation of the N slots");
            f = 1;
        }
    }
    HIP_OK(hipSetDevice(c->device));
    kspir_keys* k = new kspir_keys();
    k->ctx = c;
    k->r = r;
    k->n1 = n1;
    k->n2 = n2;
    const int nb = n1 / 2, ng = n2 - 1, np = ilog2(r);
    size_t raw_words = (size_t)nb * 6 * 2 * kN;   // largest raw upload
    if ((size_t)ng * 2 * 2 * kN > raw_words) raw_words = (size_t)ng * 2 * 2 * kN;
    if ((size_t)np * 4 * 2 * kN > raw_words) raw_words = (size_t)np * 4 * 2 * kN;
    if ((size_t)4 * 2 * kN > raw_words) raw_words = (size_t)4 * 2 * kN;
    uint64_t* raw = nullptr;
    auto cleanup = [&](int rc) {
        if (raw) hipFree(raw);
        if (rc != KSPIR_OK) kspir_keys_destroy(k);
        return rc;
    };
    hipError_t e = hipMalloc(&raw, raw_words * 8);
    if (e != hipSuccess) return cleanup(fail_hip("hipMalloc(raw keys)", e));
#define K_OK(x) do { hipError_t e2 = (x); if (e2 != hipSuccess) return cleanup(fail_hip(#x, e2)); } while (0)
    if (!p16.empty()) {
        K_OK(hipMalloc(&k->perms, p16.size() * 2));
        K_OK(hipMemcpy(k->perms, p16.data(), p16.size() * 2, hipMemcpyHostToDevice));
    }
    if (baby) {   // stored in each rotation's slot order (needs the permutation rows)
        K_OK(hipMalloc(&k->baby, (size_t)nb * 3 * 2 * 3 * kN * 4));
        K_OK(hipMemcpy(raw, baby, (size_t)nb * 6 * 2 * kN * 8, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_keys_baby, dim3(grid1d((size_t)nb * 6 * kN)), dim3(256), 0, c->stream, k->baby, raw,
                           k->perms, nb, c->T);
        K_OK(hipStreamSynchronize(c->stream));
    }
    auto upload_q = [&](const uint64_t* src, int rows, uint32_t** dst) -> hipError_t {
        hipError_t ee = hipMalloc(dst, (size_t)rows * 2 * 2 * kN * 4);
        if (ee != hipSuccess) return ee;
        ee = hipMemcpy(raw, src, (size_t)rows * 2 * kN * 8, hipMemcpyHostToDevice);
        if (ee != hipSuccess) return ee;
        hipLaunchKernelGGL(k_keys_q, dim3(grid1d((size_t)rows * 2 * kN)), dim3(256), 0, c->stream, *dst, raw, rows, c->T);
        return hipStreamSynchronize(c->stream);
    };
    if (giant && ng > 0) K_OK(upload_q(giant, ng * 2, &k->giant));
    if (packing && np > 0) K_OK(upload_q(packing, np * 4, &k->packing));
    if (rgsw) K_OK(upload_q(rgsw, 4, &k->rgsw));
    k->ptrs = KeyPtrs{k->baby, k->giant, k->packing, k->rgsw, k->perms};
    K_OK(hipMalloc(&k->dev_ptrs, sizeof(KeyPtrs)));
    K_OK(hipMemcpy(k->dev_ptrs, &k->ptrs, sizeof(KeyPtrs), hipMemcpyHostToDevice));
#undef K_OK
    *out = k;
    return cleanup(KSPIR_OK);
}

void kspir_keys_destroy(kspir_keys* k) {
    if (!k) return;
    hipSetDevice(k->ctx->device);
    drop_graphs_everywhere(k->ctx->device, k->dev_ptrs);
    // a batch table naming these keys must not be reused
    k->ctx->kp_count = 0;
    hipFree(k->baby);
    hipFree(k->giant);
    hipFree(k->packing);
    hipFree(k->rgsw);
    hipFree(k->perms);
    hipFree(k->dev_ptrs);
    delete k;
}

int kspir_answer_device(kspir_context* c, const kspir_database* db, const kspir_keys* k, const uint64_t* query_dev,
                        uint64_t* response_dev, void* stream) {
    int rc = check_pair(c, db, k);
    if (rc) return rc;
    if (!query_dev || !response_dev) return fail(KSPIR_EINVAL, "null buffer");
    HIP_OK(hipSetDevice(c->device));
    rc = carve_workspace(c, db->r, db->n1, 1);
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    return run_answer(c, db, k->dev_ptrs, 1, query_dev, response_dev, st);
}

int kspir_answer_device_noalloc(kspir_context* c, const kspir_database* db, const kspir_keys* k,
                                const uint64_t* query_dev, uint64_t* response_dev, void* stream) {
    if (!c || !db || !k) return fail(KSPIR_EINVAL, "null handle");
    const size_t need = layout_workspace(c, db->r, db->n1, 1, nullptr, true) * 4;
    if (need > kScratchBytes)
        return fail(KSPIR_EINVAL, "kspir_answer_device_noalloc: shape r=" + std::to_string(db->r) + " N1=" +
                                      std::to_string(db->n1) + " needs " + std::to_string(need) +
                                      " bytes of workspace, above its fixed scratch of " +
                                      std::to_string(kScratchBytes) + " bytes (the r 16, N1 128 shape fits)");
    int rc = check_pair(c, db, k);
    if (rc) return rc;
    if (!query_dev || !response_dev) return fail(KSPIR_EINVAL, "null buffer");
    HIP_OK(hipSetDevice(c->device));
    if (!db->words)
        return fail(KSPIR_EINVAL, "database kept as int8 limbs only (kspir_database_compact): the no-allocation answer "
                                  "runs the word pass");
    uint32_t* ptr = g_scratch_ptr[c->device];   // set when the context was created
    if (!ptr) return fail(KSPIR_EINVAL, "no scratch address for this device");
    // every buffer the kernels use is written before it is read, so stale
    // contents of an earlier call need no clearing
    layout_workspace(c, db->r, db->n1, 1, ptr, true);
    c->cur_cb = 0;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (st != c->stream) c->last_stream = st;
    // eager launches only (no graph capture), VALU database pass (the MFMA
    // pass keeps lazily allocated limbs)
    return enqueue_answer(c, db, k->dev_ptrs, 1, query_dev, response_dev, st, nullptr, nullptr, false, false);
}

int kspir_answer_batch_device(kspir_context* c, const kspir_database* db, const kspir_keys* const* keys, int32_t nq,
                              const uint64_t* queries_dev, uint64_t* responses_dev, void* stream) {
    if (!keys || nq < 1) return fail(KSPIR_EINVAL, "empty batch");
    for (int i = 0; i < nq; ++i) {
        int rc = check_pair(c, db, keys[i]);
        if (rc) return rc;
    }
    if (!queries_dev || !responses_dev) return fail(KSPIR_EINVAL, "null buffer");
    HIP_OK(hipSetDevice(c->device));
    int rc = carve_workspace(c, db->r, db->n1, nq);
    if (rc) return rc;
    rc = stage_batch_keys(c, keys, nq);
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (st != c->stream) HIP_OK(hipStreamSynchronize(c->stream));   // key table upload runs on the context stream
    return run_answer(c, db, c->kp_dev, nq, queries_dev, responses_dev, st);
}

int kspir_answer_multi_device(kspir_context* c, int32_t nseg, const kspir_database* const* dbs,
                              const int32_t* seg_queries, const kspir_keys* const* keys, const uint64_t* queries_dev,
                              uint64_t* responses_dev, void* stream) {
    std::vector<Seg> segs;
    std::vector<const void*> sig;
    int nq = 0;
    int rc = prepare_multi(c, nseg, dbs, seg_queries, keys, segs, sig, &nq);
    if (rc) return rc;
    if (!queries_dev || !responses_dev) return fail(KSPIR_EINVAL, "null buffer");
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (st != c->stream) HIP_OK(hipStreamSynchronize(c->stream));   // key table upload runs on the context stream
    return run_graph(c, nullptr, sig, c->kp_dev, nq, queries_dev, responses_dev, st,
                     [&] { return enqueue_multi(c, segs, c->kp_dev, nq, queries_dev, responses_dev, st); });
}

int kspir_answer_multi(kspir_context* c, int32_t nseg, const kspir_database* const* dbs, const int32_t* seg_queries,
                       const kspir_keys* const* keys, const uint64_t* queries, uint64_t* responses) {
    std::vector<Seg> segs;
    std::vector<const void*> sig;
    int nq = 0;
    int rc = prepare_multi(c, nseg, dbs, seg_queries, keys, segs, sig, &nq);
    if (rc) return rc;
    if (!queries || !responses) return fail(KSPIR_EINVAL, "null buffer");
    uint64_t* qd = c->q_dev + (size_t)nq * c->S.qin64;   // second half of the query area
    uint64_t* rd = c->resp_dev + (size_t)nq * c->S.resp64;
    HIP_OK(hipMemcpyAsync(qd, queries, (size_t)nq * 4 * kN * 8, hipMemcpyHostToDevice, c->stream));
    rc = kspir_answer_multi_device(c, nseg, dbs, seg_queries, keys, qd, rd, nullptr);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(responses, rd, (size_t)nq * 2 * kN * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return KSPIR_OK;
}

int kspir_time_answer_multi(kspir_context* c, int32_t nseg, const kspir_database* const* dbs,
                            const int32_t* seg_queries, const kspir_keys* const* keys, const uint64_t* queries_dev,
                            uint64_t* responses_dev, int warmup, int iters, float* ms) {
    if (!ms || iters < 1) return fail(KSPIR_EINVAL, "bad timing arguments");
    for (int i = 0; i < warmup; ++i) {
        int rc = kspir_answer_multi_device(c, nseg, dbs, seg_queries, keys, queries_dev, responses_dev, nullptr);
        if (rc) return rc;
    }
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; ++i) {
        int rc = kspir_answer_multi_device(c, nseg, dbs, seg_queries, keys, queries_dev, responses_dev, nullptr);
        if (rc) return rc;
    }
    HIP_OK(hipEventRecord(e1, c->stream));
    HIP_OK(hipEventSynchronize(e1));
    HIP_OK(hipEventElapsedTime(ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return KSPIR_OK;
}

int kspir_answer_batch(kspir_context* c, const kspir_database* db, const kspir_keys* const* keys, int32_t nq,
                       const uint64_t* queries, uint64_t* responses) {
    if (!keys || nq < 1 || !queries || !responses) return fail(KSPIR_EINVAL, "bad batch arguments");
    if (!c || !db) return fail(KSPIR_EINVAL, "null handle");
    HIP_OK(hipSetDevice(c->device));
    int rc = carve_workspace(c, db->r, db->n1, nq);
    if (rc) return rc;
    uint64_t* qd = c->q_dev + (size_t)nq * c->S.qin64;   // second half of the query area
    HIP_OK(hipMemcpyAsync(qd, queries, (size_t)nq * 4 * kN * 8, hipMemcpyHostToDevice, c->stream));
    uint64_t* rd = c->resp_dev + (size_t)nq * c->S.resp64;
    rc = kspir_answer_batch_device(c, db, keys, nq, qd, rd, nullptr);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(responses, rd, (size_t)nq * 2 * kN * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return KSPIR_OK;
}

int kspir_answer_inspect(kspir_context* c, const kspir_database* db, const kspir_keys* k, const uint64_t* query,
                         uint64_t* response, uint64_t* cipherbuf, uint64_t* firstdim, uint64_t* ext) {
    int rc = check_pair(c, db, k);
    if (rc) return rc;
    if (!query || !response) return fail(KSPIR_EINVAL, "null buffer");
    HIP_OK(hipSetDevice(c->device));
    rc = carve_workspace(c, db->r, db->n1, 1, false);
    if (rc) return rc;
    const int r = db->r;
    uint64_t* fd_dev = c->u64tmp;
    uint64_t* ext_dev = nullptr;
    if (ext) HIP_OK(hipMalloc(&ext_dev, (size_t)r * 2 * kN * 8));
    HIP_OK(hipMemcpyAsync(c->q_dev, query, 4 * kN * 8, hipMemcpyHostToDevice, c->stream));
    rc = (firstdim || ext) ? enqueue_answer(c, db, k->dev_ptrs, 1, c->q_dev, c->resp_dev, c->stream,
                                            firstdim ? fd_dev : nullptr, ext_dev, false)
                           : run_answer(c, db, k->dev_ptrs, 1, c->q_dev, c->resp_dev, c->stream);
    if (rc) { if (ext_dev) hipFree(ext_dev); return rc; }
    HIP_OK(hipMemcpyAsync(response, c->resp_dev, 2 * kN * 8, hipMemcpyDeviceToHost, c->stream));
    if (cipherbuf) {
        hipLaunchKernelGGL(k_ctw_pack, dim3(grid1d((size_t)kN * db->n1 * 2)), dim3(256), 0, c->stream, c->ctw,
                           c->u64tmp2, db->n1);
        LAUNCH_OK("k_ctw_pack");
        HIP_OK(hipMemcpyAsync(cipherbuf, c->u64tmp2, (size_t)kN * db->n1 * 2 * 8, hipMemcpyDeviceToHost, c->stream));
    }
    if (firstdim) HIP_OK(hipMemcpyAsync(firstdim, fd_dev, (size_t)r * 2 * kN * 8, hipMemcpyDeviceToHost, c->stream));
    if (ext) HIP_OK(hipMemcpyAsync(ext, ext_dev, (size_t)r * 2 * kN * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (ext_dev) HIP_OK(hipFree(ext_dev));
    return KSPIR_OK;
}

int kspir_answer(kspir_context* c, const kspir_database* db, const kspir_keys* k, const uint64_t* query,
                 uint64_t* response) {
    return kspir_answer_inspect(c, db, k, query, response, nullptr, nullptr, nullptr);
}

int kspir_matrix_vector_mul_bsgs_rns_crt_large(kspir_context* c, const kspir_database* db, const kspir_keys* k,
                                               uint64_t* query, uint64_t* result) {
    int rc0 = check_pair(c, db, k, true);
    if (rc0) return rc0;
    if (!query || !result) return fail(KSPIR_EINVAL, "null buffer");
    HIP_OK(hipSetDevice(c->device));
    int rc = carve_workspace(c, db->r, db->n1, 1, false);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(c->q_dev, query, 4 * kN * 8, hipMemcpyHostToDevice, c->stream));
    rc = enqueue_answer(c, db, k->dev_ptrs, 1, c->q_dev, nullptr, c->stream, c->u64tmp, nullptr, true);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ctw0_to_u64, dim3(2 * kN / 256), dim3(256), 0, c->stream, c->ctw, c->resp_dev, db->n1, c->T);
    LAUNCH_OK("k_ctw0_to_u64");
    HIP_OK(hipMemcpyAsync(result, c->u64tmp, (size_t)db->r * 2 * kN * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(query, c->resp_dev, 2 * kN * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return KSPIR_OK;
}

int kspir_external_product(kspir_context* c, const kspir_keys* k, uint64_t* result, uint64_t* first_dim,
                           int first_dim_is_ntt) {
    if (!c || !k || !result || !first_dim) return fail(KSPIR_EINVAL, "null argument");
    if (!k->rgsw) return fail(KSPIR_EINVAL, "keys lack the RGSW ciphertext");
    HIP_OK(hipSetDevice(c->device));
    int rc = carve_workspace(c, 1, k->n1, 1);
    if (rc) return rc;
    hipStream_t st = c->stream;
    const Tabs& T = c->T;
    const QS& S = c->S;
    uint64_t* io = c->u64tmp;   // [0]: first_dim, [1]: result words
    HIP_OK(hipMemcpyAsync(io, first_dim, 2 * kN * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(io + 2 * kN, result, 2 * kN * 8, hipMemcpyHostToDevice, st));
    if (first_dim_is_ntt) {
        hipLaunchKernelGGL(k_u64_to_rns, dim3(1, 2, 2), dim3(kNttThreads), 0, st, io, c->fd, 0, T);
        LAUNCH_OK("k_u64_to_rns");
        hipLaunchKernelGGL(k_response, dim3(2, 1), dim3(kNttThreads), 0, st, c->fd, io, (size_t)4 * kN, T);
        LAUNCH_OK("k_response");
    }
    hipLaunchKernelGGL(k_u64_to_rns, dim3(1, 2, 2), dim3(kNttThreads), 0, st, io, c->ecoef, 0, T);
    LAUNCH_OK("k_u64_to_rns");
    hipLaunchKernelGGL(k_ext_digit, dim3(1, 4, 2), dim3(kNttThreads), 0, st, c->ecoef, k->dev_ptrs, c->epart, S, T);
    LAUNCH_OK("k_ext_digit");
    hipLaunchKernelGGL(k_ext_sum, dim3(grid1d(4 * kN)), dim3(256), 0, st, c->epart, c->extA, 1, 1, S, T);
    LAUNCH_OK("k_ext_sum");
    // result words are NTT-domain accumulators: add them
    hipLaunchKernelGGL(k_u64_to_rns, dim3(1, 2, 2), dim3(kNttThreads), 0, st, io + 2 * kN, c->extB, 0, T);
    LAUNCH_OK("k_u64_to_rns");
    hipLaunchKernelGGL(k_add_planes, dim3(grid1d(4 * kN)), dim3(256), 0, st, c->extA, c->extB, 1, T);
    LAUNCH_OK("k_add_planes");
    hipLaunchKernelGGL(k_response, dim3(2, 1), dim3(kNttThreads), 0, st, c->extA, io + 2 * kN, (size_t)4 * kN, T);
    LAUNCH_OK("k_response");
    HIP_OK(hipMemcpyAsync(result, io + 2 * kN, 2 * kN * 8, hipMemcpyDeviceToHost, st));
    if (first_dim_is_ntt) HIP_OK(hipMemcpyAsync(first_dim, io, 2 * kN * 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return KSPIR_OK;
}

int kspir_packing_rlwes(kspir_context* c, const kspir_keys* k, uint64_t* result, uint64_t* rlwes, int rlwes_are_ntt) {
    if (!c || !k || !result || !rlwes) return fail(KSPIR_EINVAL, "null argument");
    const int r = k->r;
    if (r > 1 && !k->packing) return fail(KSPIR_EINVAL, "keys lack packing keys");
    HIP_OK(hipSetDevice(c->device));
    int rc = carve_workspace(c, r, k->n1, 1);
    if (rc) return rc;
    hipStream_t st = c->stream;
    const Tabs& T = c->T;
    HIP_OK(hipMemcpyAsync(c->u64tmp, rlwes, (size_t)r * 2 * kN * 8, hipMemcpyHostToDevice, st));
    // the packing tree works on NTT-domain planes
    hipLaunchKernelGGL(k_u64_to_rns, dim3(r, 2, 2), dim3(kNttThreads), 0, st, c->u64tmp, c->extA,
                       rlwes_are_ntt ? 0 : 1, T);
    LAUNCH_OK("k_u64_to_rns");
    uint32_t* packed = nullptr;
    rc = enqueue_packing(c, k->dev_ptrs, r, 1, st, &packed);
    if (rc) return rc;
    hipLaunchKernelGGL(k_response, dim3(2, 1), dim3(kNttThreads), 0, st, packed, c->resp_dev, c->S.ext, T);
    LAUNCH_OK("k_response");
    HIP_OK(hipMemcpyAsync(result, c->resp_dev, 2 * kN * 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return KSPIR_OK;
}

// ------------------------------------------------------------ runtime plumbing

int kspir_context_set(kspir_context* c, const char* key, int64_t value) {
    if (!c || !key) return fail(KSPIR_EINVAL, "null argument");
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipStreamSynchronize(c->stream));
    drop_graphs(c);   // captured launches carry the old choices
    const std::string k(key);
    if (k == "mfma_min_batch") {
        if (value < 1) return fail(KSPIR_EINVAL, "mfma_min_batch must be >= 1");
        c->mfma_min_batch = (int)value;
    } else if (k == "mfma_max_chunk") {
        if (value < 1) return fail(KSPIR_EINVAL, "mfma_max_chunk must be >= 1");
        c->mfma_max_chunk = (int)value;
    } else if (k == "fused_min_batch") {
        if (value < 1) return fail(KSPIR_EINVAL, "fused_min_batch must be >= 1");
        c->fused_min_batch = (int)value;
    } else if (k == "fused_pack") {
        c->fused_pack = value != 0;
    } else if (k == "mfma_direct") {
        c->mfma_direct = value != 0;
    } else if (k == "ext_fused") {
        c->ext_fused = value != 0;
    } else if (k == "chunk_blocks") {
        if (value < 0 || (value & (value - 1))) return fail(KSPIR_EINVAL, "chunk_blocks must be 0 or a power of two");
        c->chunk_blocks = (int)value;
    } else if (k == "graphs") {
        c->use_graphs = value != 0;
    } else if (k == "pack_fused_min_wgs") {
        if (value < 0) return fail(KSPIR_EINVAL, "pack_fused_min_wgs must be >= 0");
        c->pack_fused_min_wgs = value;
    } else if (k == "giant_wgs_per_cu") {
        if (value != 1 && value != 2) return fail(KSPIR_EINVAL, "giant_wgs_per_cu must be 1 or 2");
        c->giant_wgs_per_cu = (int)value;
    } else if (k == "giant_dma") {
        c->giant_dma = value != 0;
    } else if (k == "db_waves") {
        if (value != 2 && value != 4 && value != 8 && value != 16)
            return fail(KSPIR_EINVAL, "db_waves must be 2, 4, 8 or 16");
        c->db_waves = (int)value;
    } else if (k == "serve_multi") {
        c->serve_multi = value != 0;
    } else if (k == "serve_graphs") {
        c->serve_graphs = value != 0;
    } else if (k == "serve_fail_batches") {
        if (value < 0) return fail(KSPIR_EINVAL, "serve_fail_batches must be >= 0");
        c->serve_fail_batches = value;
    } else {
        return fail(KSPIR_EINVAL, "unknown option " + k);
    }
    return KSPIR_OK;
}

int kspir_context_get(kspir_context* c, const char* key, int64_t* value) {
    if (!c || !key || !value) return fail(KSPIR_EINVAL, "null argument");
    const std::string k(key);
    if (k == "mfma_min_batch") *value = c->mfma_min_batch;
    else if (k == "mfma_max_chunk") *value = c->mfma_max_chunk;
    else if (k == "fused_min_batch") *value = c->fused_min_batch;
    else if (k == "fused_pack") *value = c->fused_pack;
    else if (k == "mfma_direct") *value = c->mfma_direct;
    else if (k == "ext_fused") *value = c->ext_fused;
    else if (k == "chunk_blocks") *value = c->chunk_blocks;
    else if (k == "graphs") *value = c->use_graphs;
    else if (k == "db_waves") *value = c->db_waves;
    else if (k == "serve_multi") *value = c->serve_multi;
    else if (k == "serve_graphs") *value = c->serve_graphs;
    else if (k == "serve_fail_batches") *value = c->serve_fail_batches;
    else if (k == "giant_wgs_per_cu") *value = c->giant_wgs_per_cu;
    else if (k == "giant_dma") *value = c->giant_dma;
    else if (k == "pack_fused_min_wgs") *value = c->pack_fused_min_wgs;
    else return fail(KSPIR_EINVAL, "unknown option " + k);
    return KSPIR_OK;
}

int kspir_device_alloc(kspir_context* c, size_t bytes, void** out) {
    if (!c || !out) return fail(KSPIR_EINVAL, "null argument");
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipMalloc(out, bytes));
    return KSPIR_OK;
}
int kspir_device_free(kspir_context* c, void* p) {
    if (!c) return fail(KSPIR_EINVAL, "null context");
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipFree(p));
    return KSPIR_OK;
}
int kspir_memcpy_h2d(kspir_context* c, void* dst, const void* src, size_t bytes) {
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return KSPIR_OK;
}
int kspir_memcpy_d2h(kspir_context* c, void* dst, const void* src, size_t bytes) {
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return KSPIR_OK;
}
int kspir_synchronize(kspir_context* c) {
    HIP_OK(hipStreamSynchronize(c->stream));
    return KSPIR_OK;
}

int kspir_time_answer_batch(kspir_context* c, const kspir_database* db, const kspir_keys* const* keys, int32_t nq,
                            const uint64_t* queries_dev, uint64_t* responses_dev, int warmup, int iters, float* ms) {
    if (!ms || iters < 1) return fail(KSPIR_EINVAL, "bad timing arguments");
    for (int i = 0; i < warmup; ++i) {
        int rc = nq == 1 ? kspir_answer_device(c, db, keys[0], queries_dev, responses_dev, nullptr)
                         : kspir_answer_batch_device(c, db, keys, nq, queries_dev, responses_dev, nullptr);
        if (rc) return rc;
    }
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; ++i) {
        int rc = nq == 1 ? kspir_answer_device(c, db, keys[0], queries_dev, responses_dev, nullptr)
                         : kspir_answer_batch_device(c, db, keys, nq, queries_dev, responses_dev, nullptr);
        if (rc) return rc;
    }
    HIP_OK(hipEventRecord(e1, c->stream));
    HIP_OK(hipEventSynchronize(e1));
    HIP_OK(hipEventElapsedTime(ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return KSPIR_OK;
}

int kspir_time_answer(kspir_context* c, const kspir_database* db, const kspir_keys* k, const uint64_t* query_dev,
                      uint64_t* response_dev, int warmup, int iters, float* ms) {
    return kspir_time_answer_batch(c, db, &k, 1, query_dev, response_dev, warmup, iters, ms);
}

int kspir_time_db_pass(kspir_context* c, const kspir_database* db, const kspir_keys* k, const uint64_t* query_dev,
                       int iters, float* ms_per_launch) {
    if (!ms_per_launch || iters < 1) return fail(KSPIR_EINVAL, "bad timing arguments");
    int rc = check_pair(c, db, k, true);
    if (rc) return rc;
    if (!db->words)
        return fail(KSPIR_EINVAL, "database kept as int8 limbs only (kspir_database_compact): no word pass to time");
    HIP_OK(hipSetDevice(c->device));
    rc = carve_workspace(c, db->r, db->n1, 1, false);
    if (rc) return rc;
    rc = enqueue_answer(c, db, k->dev_ptrs, 1, query_dev, c->resp_dev, c->stream, nullptr, nullptr, true);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; ++i) {
        rc = launch_db_product(c, db, 1, c->stream);
        if (rc) return rc;
    }
    HIP_OK(hipEventRecord(e1, c->stream));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_launch = ms / iters;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return KSPIR_OK;
}

int kspir_time_db_pass_batch(kspir_context* c, const kspir_database* db, const kspir_keys* const* keys, int32_t nq,
                             const uint64_t* queries_dev, int iters, float* ms3) {
    if (!ms3 || iters < 1 || !keys || nq < 2) return fail(KSPIR_EINVAL, "bad timing arguments (a batch needs 2 queries)");
    for (int i = 0; i < nq; ++i) {
        int rc = check_pair(c, db, keys[i], true);
        if (rc) return rc;
    }
    if (!mfma_shape(db)) return fail(KSPIR_EINVAL, "database shape has no MFMA pass");
    HIP_OK(hipSetDevice(c->device));
    int rc = carve_workspace(c, db->r, db->n1, nq, false);
    if (rc) return rc;
    rc = stage_batch_keys(c, keys, nq);
    if (rc) return rc;
    // one first dimension fills the rotation words and makes the lazy allocations
    rc = enqueue_answer(c, db, c->kp_dev, nq, queries_dev, nullptr, c->stream, nullptr, nullptr, true);
    if (rc) return rc;
    rc = launch_db_batched(c, db, 0, nq, c->stream, 0);   // makes its lazy allocations, launches nothing
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    const int part[3] = {1, 2, 4};
    // the direct pass in the slot-group-major layout, as the batched pipeline runs it
    const bool zmaj = c->mfma_direct != 0;
    for (int t = 0; t < 3; ++t) {
        HIP_OK(hipEventRecord(e0, c->stream));
        for (int i = 0; i < iters; ++i) {
            rc = launch_db_batched(c, db, 0, nq, c->stream, part[t], 0, 0, zmaj);
            if (rc) return rc;
        }
        HIP_OK(hipEventRecord(e1, c->stream));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        ms3[t] = ms / iters;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return KSPIR_OK;
}

int kspir_time_copy(kspir_context* c, void* dst, const void* src, size_t bytes, int iters, float* ms_per_launch) {
    if (!c || !dst || !src || !ms_per_launch || iters < 1 || bytes == 0 || bytes % 16)
        return fail(KSPIR_EINVAL, "bad copy arguments");
    HIP_OK(hipSetDevice(c->device));
    const size_t n = bytes / 16;
    const int grid = 256 * 8;
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; ++i) {
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, (const uint4*)src, (uint4*)dst, n);
        LAUNCH_OK("k_copy");
    }
    HIP_OK(hipEventRecord(e1, c->stream));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_launch = ms / iters;
    HIP_OK(hipEventDestroy(e0));
    HIP_OK(hipEventDestroy(e1));
    return KSPIR_OK;
}

}  // extern "C"

// ------------------------------------------------------------ serving loop

struct kspir_server {
    struct Req {
        uint64_t id;
        const kspir_database* db;
        const kspir_keys* keys;
        const uint64_t* query;
        uint64_t* response;
        std::chrono::steady_clock::time_point t0;
    };
    struct Done {
        int rc;
        double ms;
    };
    // one lane per worker thread: its own context (stream, workspace,
    // graphs) and pinned staging, so batches of different lanes overlap on
    // the device; lane 0 drives the caller's context
    struct Lane {
        kspir_context* ctx = nullptr;
        bool owned = false;
        uint64_t* qh = nullptr;   // pinned staging, max_batch queries / responses
        uint64_t* rh = nullptr;
        std::thread worker;
    };
    kspir_context* ctx = nullptr;
    int max_batch = 64;
    size_t budget = 0;            // per lane
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req> pending;
    std::unordered_map<uint64_t, Done> done;
    std::unordered_set<uint64_t> live;   // submitted and not yet collected by a wait
    uint64_t next_id = 1, batches = 0, answered = 0, spanning = 0;   // spanning: batches over several databases
    bool stop = false;
    std::vector<Lane> lanes;

    // The next batch, taken from `pending` (lock held): the database of the
    // oldest waiting query and as many of its queries, in arrival order, as
    // the batch cap and the lane's workspace budget admit.  With the lane's
    // option serve_multi, waiting queries of other databases of the same N1
    // and at most as many blocks follow in arrival order (at most kMaxSeg
    // databases per batch), each admitted while the batch's footprint at its
    // total count fits the budget; such a batch is answered by one launch per
    // stage (kspir_answer_multi).  Only smaller databases join: the batch
    // then takes about as long as the first database's queries alone would,
    // while a query of a larger database would make every query of the batch
    // wait for its longer giant steps.  take is grouped by database, in the
    // order of dbs / seg_n.
    void pick(Lane& L, std::vector<Req>& take, std::vector<const kspir_database*>& dbs, std::vector<int32_t>& seg_n) {
        take.clear();
        dbs.clear();
        seg_n.clear();
        const kspir_database* db = pending.front().db;
        // the largest batch whose device footprint fits the lane's budget
        size_t cap = (size_t)max_batch;
        while (cap > 1 && batch_footprint_bytes(L.ctx, db->r, db->n1, (int)cap) > budget) --cap;
        for (auto it = pending.begin(); it != pending.end() && take.size() < cap;) {
            if (it->db == db) {
                take.push_back(*it);
                it = pending.erase(it);
            } else {
                ++it;
            }
        }
        // the footprint is not monotonic in the batch size (a few queries of
        // a small database take the unchunked path, with every block's rows
        // per query): the batch actually taken must fit too
        while (take.size() > 1 && batch_footprint_bytes(L.ctx, db->r, db->n1, (int)take.size()) > budget) {
            pending.push_front(take.back());
            take.pop_back();
        }
        dbs.push_back(db);
        seg_n.push_back((int32_t)take.size());
        if (!L.ctx->serve_multi || !multi_options(L.ctx) || !mfma_shape(db)) return;
        std::vector<std::vector<Req>> more;
        int total = (int)take.size();
        for (auto it = pending.begin(); it != pending.end() && total < max_batch;) {
            const kspir_database* d = it->db;
            size_t k = 0;
            while (k < dbs.size() && dbs[k] != d) ++k;
            const bool fits = d->n1 == db->n1 && d->r <= db->r && mfma_shape(d) &&
                              (k < dbs.size() || (int)dbs.size() < kMaxSeg) &&
                              multi_footprint_bytes(L.ctx, db->r, db->n1, total + 1) <= budget;
            // the first database's queries left behind were cut by the cap
            if (!fits || k == 0) {
                ++it;
                continue;
            }
            if (k == dbs.size()) {
                dbs.push_back(d);
                seg_n.push_back(0);
                more.emplace_back();
            }
            more[k - 1].push_back(*it);
            ++seg_n[k];
            ++total;
            it = pending.erase(it);
        }
        for (auto& m : more) take.insert(take.end(), m.begin(), m.end());
    }

    void run(Lane& L) {
        std::vector<Req> take;
        std::vector<const kspir_keys*> keys;
        std::vector<const kspir_database*> dbs;
        std::vector<int32_t> seg_n;
        for (;;) {
            bool more = false;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv_work.wait(lock, [&] { return stop || !pending.empty(); });
                if (pending.empty()) return;   // stop requested and drained
                pick(L, take, dbs, seg_n);
                more = !pending.empty();
            }
            if (more) cv_work.notify_one();   // another lane may take the rest
            const int nq = (int)take.size();
            keys.resize(nq);
            for (int i = 0; i < nq; ++i) {
                std::memcpy(L.qh + (size_t)i * 4 * kN, take[i].query, 4 * kN * 8);
                keys[i] = take[i].keys;
            }
            // batch sizes vary from batch to batch, so that most batches
            // would be a new capture: without serve_graphs each batch is
            // launched eagerly (the two measure alike at 4 and 16 GB)
            const int graphs = L.ctx->use_graphs;
            L.ctx->use_graphs = graphs && L.ctx->serve_graphs;
            int rc;
            if (L.ctx->serve_fail_batches > 0) {   // a failure forced from the host: nothing is launched
                --L.ctx->serve_fail_batches;
                rc = KSPIR_EINJECTED;
            } else {
                rc = dbs.size() == 1 ? kspir_answer_batch(L.ctx, dbs[0], keys.data(), nq, L.qh, L.rh)
                                     : kspir_answer_multi(L.ctx, (int32_t)dbs.size(), dbs.data(), seg_n.data(),
                                                          keys.data(), L.qh, L.rh);
            }
            L.ctx->use_graphs = graphs;
            const auto t1 = std::chrono::steady_clock::now();
            {
                std::lock_guard<std::mutex> lock(mu);
                for (int i = 0; i < nq; ++i) {
                    if (rc == KSPIR_OK) std::memcpy(take[i].response, L.rh + (size_t)i * 2 * kN, 2 * kN * 8);
                    done[take[i].id] = Done{rc, std::chrono::duration<double, std::milli>(t1 - take[i].t0).count()};
                }
                ++batches;
                if (dbs.size() > 1) ++spanning;
                answered += (uint64_t)nq;
            }
            cv_done.notify_all();
        }
    }
};

extern "C" {

int kspir_client_draws(int32_t r, int32_t n1, uint64_t* draws) {
    if (!draws || !pow2(r) || r > kN || !pow2(n1) || n1 < 2 || n1 > kN / 2)
        return fail(KSPIR_EINVAL, "r must be a power of two <= N, N1 a power of two in [2, N/2]");
    const int n2 = kN / 2 / n1;
    *draws = (uint64_t)kN * (18 + 15 * (n1 / 2) + 6 * (n2 - 1) + 12 * ilog2(r));
    return KSPIR_OK;
}

}  // extern "C"

namespace {

// odd x: x^(N-1) is its inverse mod 2N (the units mod 2N form a group of order N)
uint32_t inv2n(uint32_t x) {
    uint32_t r = 1, b = x & (2 * kN - 1);
    for (uint32_t e = kN - 1; e; e >>= 1, b = (b * b) & (2 * kN - 1))
        if (e & 1) r = (r * b) & (2 * kN - 1);
    return r;
}

uint32_t pow2mod(uint32_t k, uint64_t q) {
    uint64_t v = 1;
    for (uint32_t i = 0; i < k; ++i) v = (v << 1) % q;
    return (uint32_t)v;
}

uint32_t pow5(uint64_t e) {   // 5^e mod 2N
    uint32_t v = 1;
    for (uint64_t i = 0; i < e; ++i) v = v * 5 % (2 * kN);
    return v;
}

// Encryption jobs of one client, draws counted from `d` (its first draw),
// in the reference's order.  Outputs are device pointers in the reference
// layouts; each helper returns the draw count it used.
struct JobSink {
    std::vector<kspir::ClientJob>& jobs;
    int ci;
    void add(int32_t type, uint64_t draw, int32_t msg, uint32_t mfexp, uint32_t num, uint64_t* a, uint64_t* b,
             uint64_t* a2, uint64_t* b2) {
        kspir::ClientJob J{};
        J.type = type;
        J.client = ci;
        J.draw = (uint32_t)draw;
        J.msg = msg;
        J.msg_inv = msg >= 1 ? inv2n((uint32_t)msg) : 0;
        J.mf[0] = pow2mod(mfexp, kQ1);
        J.mf[1] = pow2mod(mfexp, kQ2);
        J.mf_bs = pow2mod(mfexp, kBs);
        J.num = num;
        J.a = a; J.b = b; J.a2 = a2; J.b2 = b2;
        jobs.push_back(J);
    }
    // RGSWCiphertext::keyGen(secret, index), lwe.cpp:297-352: gadget 2^20,
    // 2^38; per level the positive row 2+i, then the negated row i
    uint64_t rgsw(uint64_t d, int32_t index, uint64_t* out, bool reverse = true) {
        const uint64_t N = kN;
        for (int i = 0; i < 2; ++i) {
            add(kspir::kJobRlwe, d + 6 * N * i, -index, 20 + 18 * i, 1, out + (2 + i) * 2 * N,
                out + (2 + i) * 2 * N + N, nullptr, nullptr);
            jobs.back().msg_plus = !reverse;
            add(kspir::kJobRgswNeg, d + 6 * N * i + 3 * N, -index, 20 + 18 * i, 1, out + i * 2 * N,
                out + i * 2 * N + N, nullptr, nullptr);
            jobs.back().msg_plus = !reverse;
        }
        return 12 * N;
    }
    // bsgsKeyGen (lwe.cpp:1143-1160): baby keys (lwe.cpp:1025-1070, index
    // 5^(k+1), gadget 2^20, 2^40, 2^60) then giant keys (lwe.cpp:1072-1101,
    // index 5^(N1*(k+1)), gadget 2^20, 2^38)
    uint64_t bsgs(uint64_t d, int n1, uint64_t* baby, uint64_t* giant) {
        const uint64_t N = kN;
        const int nb = n1 / 2, n2 = kN / 2 / n1;
        for (int k = 0; k < nb; ++k) {
            const uint32_t idx = pow5(k + 1);
            for (int i = 0; i < 3; ++i) {
                uint64_t* kq = baby + ((size_t)k * 6 + i) * 2 * N;
                uint64_t* kb = baby + ((size_t)k * 6 + 3 + i) * 2 * N;
                add(kspir::kJobBaby, d + 15 * N * k + 5 * N * i, (int32_t)idx, 20 * (i + 1), 1, kq, kq + N, kb,
                    kb + N);
            }
        }
        const uint64_t g0 = d + 15 * N * nb;
        for (int k = 0; k < n2 - 1; ++k) {
            const uint32_t idx = pow5((uint64_t)n1 * (k + 1));
            for (int i = 0; i < 2; ++i) {
                uint64_t* o = giant + ((size_t)k * 2 + i) * 2 * N;
                add(kspir::kJobRlwe, g0 + 6 * N * k + 3 * N * i, (int32_t)idx, 20 + 18 * i, 1, o, o + N, nullptr,
                    nullptr);
            }
        }
        return 15 * N * nb + 6 * N * (n2 - 1);
    }
    // AutoKey::keyGen(secret, num, packing) (lwe.cpp:472-550): level li,
    // index N/2^li + 1, gadget 1, 2^14, 2^28, 2^42, error times num
    uint64_t packing(uint64_t d, int num, uint64_t* out, bool rlwe = true) {
        const uint64_t N = kN;
        std::vector<int32_t> idx;
        if (rlwe)
            for (int i = 1; i < num; i <<= 1) idx.push_back(kN / i + 1);
        else
            for (int i = 2; i <= num + 1; i <<= 1) idx.push_back(i + 1);
        for (size_t li = 0; li < idx.size(); ++li)
            for (int j = 0; j < 4; ++j) {
                uint64_t* o = out + (li * 4 + j) * 2 * N;
                add(kspir::kJobRlwe, d + 12 * N * li + 3 * N * j, idx[li], 14 * j, (uint32_t)num, o, o + N, nullptr,
                    nullptr);
            }
        return 12 * N * idx.size();
    }
};

// compute_query_encode (utils.cpp:1657-1675): the two NTT slots of target column `row`
void query_slots(int row, int32_t* out2) {
    static std::vector<int32_t> slot_of;
    static std::vector<uint32_t> rot;
    if (slot_of.empty()) {
        slot_of.assign(kN, 0);
        for (int s = 0; s < kN; ++s) slot_of[(2 * bitrev12(s) + 1) >> 1] = s;
        rot.assign(kN, 0);
        rot[0] = 1;
        rot[kN - 1] = 2 * kN - 1;
        for (int i = 1; i < kN / 2; ++i) {
            rot[i] = rot[i - 1] * 5 % (2 * kN);
            rot[kN - 1 - i] = rot[kN - i] * 5 % (2 * kN);
        }
    }
    out2[0] = slot_of[rot[row] >> 1];
    out2[1] = slot_of[rot[kN - 1 - row] >> 1];
}

kspir::ClientConsts client_consts() {
    kspir::ClientConsts K{};
    const uint32_t qs[3] = {kQ1, kQ2, kBs};
    for (int p = 0; p < 3; ++p) K.r2[p] = (uint32_t)(((u128)1 << 64) % qs[p]);
    K.bs_mod_q[0] = kBs % kQ1;
    K.bs_mod_q[1] = kBs % kQ2;
    return K;
}

int client_tables(kspir_context* c) {
    if (c->rng_jump) return KSPIR_OK;
    const kspir_rng::Mat jm = kspir_rng::power(kspir::kRngChunk);
    HIP_OK(hipMalloc(&c->rng_jump, jm.size() * 4));
    HIP_OK(hipMemcpy(c->rng_jump, jm.data(), jm.size() * 4, hipMemcpyHostToDevice));
    // twosteps.cpp:305: round((long double)m * crtMod / p) in the host's
    // long double, for every plaintext coefficient m < p
    std::vector<uint64_t> enc(kP);
    for (uint32_t m = 0; m < kP; ++m) enc[m] = (uint64_t)roundl((long double)m * kQ / kP);
    HIP_OK(hipMalloc(&c->enc_tab, enc.size() * 8));
    HIP_OK(hipMemcpy(c->enc_tab, enc.data(), enc.size() * 8, hipMemcpyHostToDevice));
    return KSPIR_OK;
}

// device allocations released together
struct DevScope {
    hipStream_t st;
    std::vector<void*> mem;
    hipError_t alloc(void** p, size_t bytes) {
        hipError_t e = hipMalloc(p, bytes ? bytes : 8);
        if (e == hipSuccess) mem.push_back(*p);
        return e;
    }
    ~DevScope() {
        hipStreamSynchronize(st);
        for (void* p : mem) hipFree(p);
    }
};

}  // namespace

extern "C" {

int kspir_client_generate(kspir_context* c, int32_t r, int32_t n1, int32_t nc, const uint32_t* seeds,
                          const int32_t* target_col, const int32_t* target_packing, uint64_t* secrets,
                          uint64_t* queries, uint64_t* rgsw, uint64_t* baby, uint64_t* giant, uint64_t* packing) {
    if (!c || nc < 1 || !seeds || !target_col || !target_packing) return fail(KSPIR_EINVAL, "null argument");
    uint64_t draws = 0;
    int rc = kspir_client_draws(r, n1, &draws);
    if (rc) return rc;
    const int n2 = kN / 2 / n1, nb = n1 / 2, lr = ilog2(r);
    for (int i = 0; i < nc; ++i)
        if (target_col[i] < 0 || target_col[i] >= kN / 2 || target_packing[i] < 0 || target_packing[i] >= r)
            return fail(KSPIR_EINVAL, "target column must be < N/2 and target packing < r");
    HIP_OK(hipSetDevice(c->device));
    if ((rc = client_tables(c))) return rc;
    hipStream_t st = c->stream;
    const int nchunks = (int)((draws + kspir::kRngChunk - 1) / kspir::kRngChunk);
    const size_t stride = (size_t)nchunks * kspir::kRngChunk;
    std::vector<uint32_t> seedw((size_t)nc * kspir_rng::kLag);
    std::vector<int32_t> slots(2 * (size_t)nc);
    for (int i = 0; i < nc; ++i) {
        kspir_rng::seed_window(seeds[i], &seedw[(size_t)i * kspir_rng::kLag]);
        query_slots(target_col[i], &slots[2 * (size_t)i]);
    }
    const size_t N = kN;
    const size_t n_sec = (size_t)nc * N, n_q = (size_t)nc * 4 * N, n_rg = (size_t)nc * 8 * N;
    const size_t n_baby = (size_t)nc * nb * 12 * N, n_giant = (size_t)nc * (n2 - 1) * 4 * N;
    const size_t n_pack = (size_t)nc * lr * 8 * N;
    DevScope ds{st, {}};
#define C_OK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return fail_hip(#x, e_); } while (0)
    uint32_t *d_seedw, *d_win, *d_stream, *d_sm;
    int32_t *d_scoef, *d_slots;
    uint64_t *d_sec, *d_q, *d_rg, *d_baby, *d_giant, *d_pack;
    kspir::ClientJob* d_jobs;
    C_OK(ds.alloc((void**)&d_seedw, seedw.size() * 4));
    C_OK(ds.alloc((void**)&d_win, (size_t)nc * nchunks * kspir_rng::kLag * 4));
    C_OK(ds.alloc((void**)&d_stream, (size_t)nc * stride * 4));
    C_OK(ds.alloc((void**)&d_scoef, n_sec * 4));
    C_OK(ds.alloc((void**)&d_sm, n_sec * 3 * 4));
    C_OK(ds.alloc((void**)&d_slots, slots.size() * 4));
    C_OK(ds.alloc((void**)&d_sec, n_sec * 8));
    C_OK(ds.alloc((void**)&d_q, n_q * 8));
    C_OK(ds.alloc((void**)&d_rg, n_rg * 8));
    C_OK(ds.alloc((void**)&d_baby, n_baby * 8));
    C_OK(ds.alloc((void**)&d_giant, n_giant * 8));
    C_OK(ds.alloc((void**)&d_pack, n_pack * 8));
    // draws of one client after srand: secret [0, N), query [N, 6N), then
    // RGSW, BSGS and packing keys (test-pir.cpp:20-24, 68-89)
    std::vector<kspir::ClientJob> jobs;
    for (int ci = 0; ci < nc; ++ci) {
        JobSink js{jobs, ci};
        uint64_t d = 6 * N;
        d += js.rgsw(d, target_packing[ci], d_rg + (size_t)ci * 8 * N);
        d += js.bsgs(d, n1, d_baby + (size_t)ci * nb * 12 * N, d_giant + (size_t)ci * (n2 - 1) * 4 * N);
        d += js.packing(d, r, d_pack + (size_t)ci * lr * 8 * N);
        if (d != draws) return fail(KSPIR_EHIP, "client draw schedule does not add up");
    }
    C_OK(ds.alloc((void**)&d_jobs, jobs.size() * sizeof(kspir::ClientJob)));
    C_OK(hipMemcpyAsync(d_seedw, seedw.data(), seedw.size() * 4, hipMemcpyHostToDevice, st));
    C_OK(hipMemcpyAsync(d_slots, slots.data(), slots.size() * 4, hipMemcpyHostToDevice, st));
    C_OK(hipMemcpyAsync(d_jobs, jobs.data(), jobs.size() * sizeof(kspir::ClientJob), hipMemcpyHostToDevice, st));
    const kspir::ClientConsts K = client_consts();
    const Tabs& T = c->T;
    hipLaunchKernelGGL(kspir::k_rng_windows, dim3(nc), dim3(64), 0, st, d_seedw, c->rng_jump, d_win, nchunks);
    LAUNCH_OK("k_rng_windows");
    hipLaunchKernelGGL(kspir::k_rng_stream, dim3((nc * nchunks + 63) / 64), dim3(64), 0, st, d_win, d_stream, nc,
                       nchunks);
    LAUNCH_OK("k_rng_stream");
    hipLaunchKernelGGL(kspir::k_client_secret, dim3(nc), dim3(kNttThreads), 0, st, d_stream, stride, d_scoef, d_sm, K,
                       T);
    LAUNCH_OK("k_client_secret");
    hipLaunchKernelGGL(kspir::k_client_query, dim3(nc), dim3(kNttThreads), 0, st, d_stream, stride, (size_t)N, d_sm,
                       d_slots, c->enc_tab, d_q, K, T);
    LAUNCH_OK("k_client_query");
    hipLaunchKernelGGL(kspir::k_client_rlwe, dim3((unsigned)jobs.size()), dim3(kNttThreads), 0, st, d_jobs, d_stream,
                       stride, d_scoef, d_sm, K, T);
    LAUNCH_OK("k_client_rlwe");
    hipLaunchKernelGGL(kspir::k_client_secret_words, dim3(grid1d(n_sec)), dim3(256), 0, st, d_scoef, d_sec, n_sec);
    LAUNCH_OK("k_client_secret_words");
    if (secrets) C_OK(hipMemcpyAsync(secrets, d_sec, n_sec * 8, hipMemcpyDeviceToHost, st));
    if (queries) C_OK(hipMemcpyAsync(queries, d_q, n_q * 8, hipMemcpyDeviceToHost, st));
    if (rgsw) C_OK(hipMemcpyAsync(rgsw, d_rg, n_rg * 8, hipMemcpyDeviceToHost, st));
    if (baby) C_OK(hipMemcpyAsync(baby, d_baby, n_baby * 8, hipMemcpyDeviceToHost, st));
    if (giant && n_giant) C_OK(hipMemcpyAsync(giant, d_giant, n_giant * 8, hipMemcpyDeviceToHost, st));
    if (packing && n_pack) C_OK(hipMemcpyAsync(packing, d_pack, n_pack * 8, hipMemcpyDeviceToHost, st));
    C_OK(hipStreamSynchronize(st));
    return KSPIR_OK;
}

int kspir_client_encrypt(kspir_context* c, int32_t kind, const uint64_t* secret, int64_t param, int32_t n1,
                         const uint32_t* draws, uint64_t ndraws, uint64_t* out) {
    if (!c || !secret || !draws || !out) return fail(KSPIR_EINVAL, "null argument");
    const uint64_t N = kN;
    std::vector<int32_t> sc(kN);
    for (int i = 0; i < kN; ++i) {
        if (secret[i] == 0) sc[i] = 0;
        else if (secret[i] == 1) sc[i] = 1;
        else if (secret[i] == kQ - 1) sc[i] = -1;
        else return fail(KSPIR_EINVAL, "secret coefficients must be 0, 1 or crtMod - 1");
    }
    uint64_t need = 0;
    size_t out_words = 0;
    int n2 = 1;
    switch (kind) {
        case KSPIR_CLIENT_QUERY:
            if (param < 0 || param >= kN / 2) return fail(KSPIR_EINVAL, "query row must be < N/2");
            need = 5 * N; out_words = 4 * N; break;
        case KSPIR_CLIENT_RGSW:
            if (param < 0 || param >= kN) return fail(KSPIR_EINVAL, "RGSW index must be < N");
            need = 12 * N; out_words = 8 * N; break;
        case KSPIR_CLIENT_BSGS:
            if (!pow2(n1) || n1 < 2 || n1 > kN / 2) return fail(KSPIR_EINVAL, "N1 must be a power of two in [2, N/2]");
            n2 = kN / 2 / n1;
            need = 15 * N * (n1 / 2) + 6 * N * (n2 - 1);
            out_words = (size_t)(n1 / 2) * 12 * N + (size_t)(n2 - 1) * 4 * N; break;
        case KSPIR_CLIENT_RGSW_FORWARD:
            if (param < 0 || param >= kN) return fail(KSPIR_EINVAL, "RGSW index must be < N");
            need = 12 * N; out_words = 8 * N; break;
        case KSPIR_CLIENT_PACKING:
        case KSPIR_CLIENT_PACKING_LWE: {
            if (param < 1 || param > kN) return fail(KSPIR_EINVAL, "packing num must be in [1, N]");
            int nidx = 0;
            if (kind == KSPIR_CLIENT_PACKING)
                for (int64_t i = 1; i < param; i <<= 1) ++nidx;
            else
                for (int64_t i = 2; i <= param + 1; i <<= 1) ++nidx;
            need = 12 * N * nidx;
            out_words = (size_t)nidx * 8 * N;
            break;
        }
        default:
            return fail(KSPIR_EINVAL, "unknown client operation");
    }
    if (ndraws != need) return fail(KSPIR_EINVAL, "draw count does not match the operation");
    HIP_OK(hipSetDevice(c->device));
    int rc = client_tables(c);
    if (rc) return rc;
    hipStream_t st = c->stream;
    DevScope ds{st, {}};
    uint32_t *d_stream, *d_sm;
    int32_t *d_scoef, *d_slots;
    uint64_t* d_out;
    C_OK(ds.alloc((void**)&d_stream, need * 4));
    C_OK(ds.alloc((void**)&d_scoef, N * 4));
    C_OK(ds.alloc((void**)&d_sm, 3 * N * 4));
    C_OK(ds.alloc((void**)&d_slots, 8));
    C_OK(ds.alloc((void**)&d_out, out_words * 8));
    C_OK(hipMemcpyAsync(d_stream, draws, need * 4, hipMemcpyHostToDevice, st));
    C_OK(hipMemcpyAsync(d_scoef, sc.data(), N * 4, hipMemcpyHostToDevice, st));
    const kspir::ClientConsts K = client_consts();
    const Tabs& T = c->T;
    hipLaunchKernelGGL(kspir::k_client_secret_given, dim3(1), dim3(kNttThreads), 0, st, d_scoef, d_sm, K, T);
    LAUNCH_OK("k_client_secret_given");
    if (kind == KSPIR_CLIENT_QUERY) {
        int32_t sl[2];
        query_slots((int)param, sl);
        C_OK(hipMemcpyAsync(d_slots, sl, 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(kspir::k_client_query, dim3(1), dim3(kNttThreads), 0, st, d_stream, (size_t)need,
                           (size_t)0, d_sm, d_slots, c->enc_tab, d_out, K, T);
        LAUNCH_OK("k_client_query");
    } else {
        std::vector<kspir::ClientJob> jobs;
        JobSink js{jobs, 0};
        if (kind == KSPIR_CLIENT_RGSW || kind == KSPIR_CLIENT_RGSW_FORWARD)
            js.rgsw(0, (int32_t)param, d_out, kind == KSPIR_CLIENT_RGSW);
        else if (kind == KSPIR_CLIENT_BSGS) js.bsgs(0, n1, d_out, d_out + (size_t)(n1 / 2) * 12 * N);
        else js.packing(0, (int)param, d_out, kind == KSPIR_CLIENT_PACKING);
        if (!jobs.empty()) {
            kspir::ClientJob* d_jobs;
            C_OK(ds.alloc((void**)&d_jobs, jobs.size() * sizeof(kspir::ClientJob)));
            C_OK(hipMemcpyAsync(d_jobs, jobs.data(), jobs.size() * sizeof(kspir::ClientJob), hipMemcpyHostToDevice,
                                st));
            hipLaunchKernelGGL(kspir::k_client_rlwe, dim3((unsigned)jobs.size()), dim3(kNttThreads), 0, st, d_jobs,
                               d_stream, (size_t)need, d_scoef, d_sm, K, T);
            LAUNCH_OK("k_client_rlwe");
        }
    }
    C_OK(hipMemcpyAsync(out, d_out, out_words * 8, hipMemcpyDeviceToHost, st));
    C_OK(hipStreamSynchronize(st));
#undef C_OK
    return KSPIR_OK;
}

int kspir_decrypt_bsgs_total(kspir_context* c, int32_t r, int32_t nq, const uint64_t* secrets,
                             const uint64_t* responses, uint64_t* messages) {
    if (!c || !secrets || !responses || !messages || nq < 1) return fail(KSPIR_EINVAL, "bad argument");
    if (r < 1 || !pow2(r) || r > kN) return fail(KSPIR_EINVAL, "r must be a power of two <= N");
    HIP_OK(hipSetDevice(c->device));
    // r^-1 mod q_p (Q is odd, so every power of two is invertible)
    uint2 qv[2];
    for (int p = 0; p < 2; ++p) {
        const u64 q = c->T.p[p].q;
        const uint32_t inv = (uint32_t)invmod((u64)r % q, q);
        qv[p] = make_uint2(inv, shoup(inv, (uint32_t)q));
    }
    const size_t sb = (size_t)nq * kN * 8, rb = (size_t)nq * 2 * kN * 8;
    uint64_t *ds = nullptr, *dr = nullptr, *dm = nullptr;
    hipError_t e = hipMalloc(&ds, sb);
    if (e == hipSuccess) e = hipMalloc(&dr, rb);
    if (e == hipSuccess) e = hipMalloc(&dm, sb);
    if (e == hipSuccess) e = hipMemcpyAsync(ds, secrets, sb, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, responses, rb, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_decrypt, dim3(nq), dim3(512), 0, c->stream, ds, dr, dm, qv[0], qv[1], c->T);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(messages, dm, sb, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (ds) hipFree(ds);
    if (dr) hipFree(dr);
    if (dm) hipFree(dm);
    if (e != hipSuccess) return fail_hip("decrypt_bsgs_total", e);
    return KSPIR_OK;
}

int kspir_gen_giant_keys_from_offline(kspir_context* c, int32_t n1, int32_t nc, const uint64_t* offline2,
                                      const uint64_t* auto2, uint64_t* chain, uint64_t* giant) {
    if (!c || !offline2 || !auto2 || !giant || nc < 1) return fail(KSPIR_EINVAL, "bad argument");
    if (!pow2(n1) || n1 < 2 || n1 > kN / 4) return fail(KSPIR_EINVAL, "N1 must be a power of two in [2, N/4]");
    const int n2 = kN / 2 / n1, nkeys = n2 - 1, nchain = n2 - 2;
    HIP_OK(hipSetDevice(c->device));
    OkdConst K{};
    for (int p = 0; p < 2; ++p) {
        const uint32_t q = c->T.p[p].q;
        K.auxinv[p] = (uint32_t)invmod(kAux % q, q);
        K.auxinv_sh[p] = shoup(K.auxinv[p], q);
    }
    K.qinv_aux = (uint32_t)invmod(kQ % kAux, kAux);
    K.qinv_aux_sh = shoup(K.qinv_aux, kAux);
    const u64 g = powmod(5, (u64)n1, 2 * kN);   // 5^N1 mod 2N (odd, so invertible)
    for (u64 x = 1; x < 2 * kN; x += 2)
        if (g * x % (2 * kN) == 1) K.ginv = (uint32_t)x;
    const size_t P = kN;                         // words per plane
    const size_t st_stride = (size_t)nkeys * 18 * P;
    const size_t b_off = (size_t)nc * 10 * 2 * P * 8, b_auto = (size_t)nc * 6 * 2 * P * 8;
    const size_t b_km = (size_t)nc * 30 * P * 4, b_st = (size_t)nc * st_stride * 4, b_c0 = (size_t)nc * 18 * P * 4;
    const size_t b_dig = (size_t)nc * kOkdRows * kOkdEll * 3 * P * 4, b_tb = (size_t)nc * kOkdRows * 3 * P * 4;
    const size_t b_giant = (size_t)nc * nkeys * 6 * P * 8;
    const size_t b_chain = chain && nchain > 0 ? (size_t)nc * nchain * 12 * P * 8 : 0;
    // device state and pinned output staging live on the context and grow
    // on demand, so repeated calls allocate nothing
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t parts[9] = {b_off, b_auto, b_km, b_st, b_c0, b_dig, b_tb, b_giant, b_chain};
    size_t total = 0;
    for (size_t b : parts) total += up(b);
    hipStream_t sm = c->stream;
    if (total > c->ows_bytes) {
        HIP_OK(hipStreamSynchronize(sm));
        if (c->ows) HIP_OK(hipFree(c->ows));
        c->ows = nullptr;
        c->ows_bytes = 0;
        HIP_OK(hipMalloc(&c->ows, total));
        c->ows_bytes = total;
    }
    if (b_giant + b_chain > c->ohost_bytes) {
        if (c->ohost) HIP_OK(hipHostFree(c->ohost));
        c->ohost = nullptr;
        c->ohost_bytes = 0;
        HIP_OK(hipHostMalloc(&c->ohost, b_giant + b_chain, hipHostMallocDefault));
        c->ohost_bytes = b_giant + b_chain;
    }
    char* base = (char*)c->ows;
    void* ptrs[9];
    for (int i = 0; i < 9; ++i) { ptrs[i] = parts[i] ? base : nullptr; base += up(parts[i]); }
    uint64_t *doff = (uint64_t*)ptrs[0], *dauto = (uint64_t*)ptrs[1], *dgiant = (uint64_t*)ptrs[7],
             *dchain = (uint64_t*)ptrs[8];
    uint32_t *km = (uint32_t*)ptrs[2], *st = (uint32_t*)ptrs[3], *c0 = (uint32_t*)ptrs[4], *dig = (uint32_t*)ptrs[5],
             *tb = (uint32_t*)ptrs[6];
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMemcpyAsync(doff, offline2, b_off, hipMemcpyHostToDevice, sm);
    if (e == hipSuccess) e = hipMemcpyAsync(dauto, auto2, b_auto, hipMemcpyHostToDevice, sm);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_okd_keys, dim3(kN / 256, 30, nc), dim3(256), 0, sm, doff, km, K);
        hipLaunchKernelGGL(k_okd_load, dim3(18, nc), dim3(kNttThreads), 0, sm, dauto, st, st_stride, c0, c->T);
        for (int k = 1; k < nkeys && e == hipSuccess; ++k) {
            const uint32_t* in = k == 1 ? c0 : st + (size_t)(k - 1) * 18 * P;
            const size_t in_stride = k == 1 ? 18 * P : st_stride;
            hipLaunchKernelGGL(k_okd_digits, dim3(kN / 256, kOkdRows, nc), dim3(256), 0, sm, in, in_stride, dig, tb,
                               K, c->T);
            hipLaunchKernelGGL(k_okd_ks, dim3(3, kOkdRows, nc), dim3(kNttThreads), 0, sm, dig, tb, km,
                               st + (size_t)k * 18 * P, st_stride, c->T);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_okd_modswitch, dim3(nkeys, 6, nc), dim3(512), 0, sm, st, st_stride, dgiant, nkeys, K,
                           c->T);
        if (dchain)
            hipLaunchKernelGGL(k_okd_export, dim3(kN / 256, nchain * 12, nc), dim3(256), 0, sm, st, st_stride, dchain,
                               nchain, c->T);
        e = hipGetLastError();
    }
    uint64_t* hg = (uint64_t*)c->ohost;
    if (e == hipSuccess) e = hipMemcpyAsync(hg, dgiant, b_giant, hipMemcpyDeviceToHost, sm);
    if (e == hipSuccess && dchain) e = hipMemcpyAsync(hg + b_giant / 8, dchain, b_chain, hipMemcpyDeviceToHost, sm);
    if (e == hipSuccess) e = hipStreamSynchronize(sm);
    if (e != hipSuccess) return fail_hip("gen_giant_keys_from_offline", e);
    std::memcpy(giant, hg, b_giant);
    if (dchain) std::memcpy(chain, hg + b_giant / 8, b_chain);
    return KSPIR_OK;
}

int kspir_query_footprint_bytes(int32_t r, int32_t n1, size_t* bytes) {
    if (!bytes || !pow2(r) || r > kN || !pow2(n1) || n1 < 32 || n1 > kN / 2) return fail(KSPIR_EINVAL, "bad shape");
    static kspir_context defaults;   // the default options, no device state
    *bytes = batch_footprint_bytes(&defaults, r, n1, 64) / 64;
    return KSPIR_OK;
}

int kspir_batch_footprint_bytes(const kspir_context* c, int32_t r, int32_t n1, int32_t nq, size_t* bytes) {
    if (!bytes || nq < 1 || !pow2(r) || r > kN || !pow2(n1) || n1 < 32 || n1 > kN / 2)
        return fail(KSPIR_EINVAL, "bad shape or batch size");
    static kspir_context defaults;   // the default options, no device state
    // the footprint only reads the options (no buffer is laid out)
    *bytes = batch_footprint_bytes(const_cast<kspir_context*>(c ? c : &defaults), r, n1, nq);
    return KSPIR_OK;
}

int kspir_server_create_lanes(kspir_context* c, int32_t max_batch, size_t workspace_budget_bytes, int32_t lanes,
                              kspir_server** out) {
    if (!c || !out || max_batch < 1 || lanes < 1 || lanes > 8) return fail(KSPIR_EINVAL, "bad server arguments");
    HIP_OK(hipSetDevice(c->device));
    kspir_server* s = new kspir_server();
    s->ctx = c;
    s->max_batch = max_batch;
    if (workspace_budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        hipError_t e = hipMemGetInfo(&fr, &tot);
        if (e != hipSuccess) { delete s; return fail_hip("hipMemGetInfo", e); }
        workspace_budget_bytes = fr / 2;
    }
    s->budget = workspace_budget_bytes / (size_t)lanes;
    s->lanes.resize(lanes);
    auto undo = [&](int rc) {
        for (auto& L : s->lanes) {
            if (L.qh) hipHostFree(L.qh);
            if (L.rh) hipHostFree(L.rh);
            if (L.owned) kspir_context_destroy(L.ctx);
        }
        delete s;
        return rc;
    };
    for (int i = 0; i < lanes; ++i) {
        kspir_server::Lane& L = s->lanes[i];
        if (i == 0) {
            L.ctx = c;
        } else {
            int rc = kspir_context_create(c->device, &L.ctx);
            if (rc) return undo(rc);
            L.owned = true;
            // the caller's tuned choices apply to every lane
            L.ctx->mfma_min_batch = c->mfma_min_batch;
            L.ctx->mfma_max_chunk = c->mfma_max_chunk;
            L.ctx->fused_min_batch = c->fused_min_batch;
            L.ctx->fused_pack = c->fused_pack;
            L.ctx->mfma_direct = c->mfma_direct;
            L.ctx->ext_fused = c->ext_fused;
            L.ctx->chunk_blocks = c->chunk_blocks;
            L.ctx->pack_fused_min_wgs = c->pack_fused_min_wgs;
            L.ctx->giant_wgs_per_cu = c->giant_wgs_per_cu;
            L.ctx->giant_dma = c->giant_dma;
            L.ctx->use_graphs = c->use_graphs;
            L.ctx->db_waves = c->db_waves;
            L.ctx->serve_multi = c->serve_multi;
            L.ctx->serve_graphs = c->serve_graphs;
            L.ctx->serve_fail_batches = c->serve_fail_batches;
        }
        hipError_t e = hipHostMalloc(&L.qh, (size_t)max_batch * 4 * kN * 8, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc(&L.rh, (size_t)max_batch * 2 * kN * 8, hipHostMallocDefault);
        if (e != hipSuccess) return undo(fail_hip("hipHostMalloc(server staging)", e));
    }
    for (auto& L : s->lanes) {
        kspir_server::Lane* lp = &L;
        L.worker = std::thread([s, lp] {
            hipSetDevice(lp->ctx->device);
            s->run(*lp);
        });
    }
    *out = s;
    return KSPIR_OK;
}

int kspir_server_create(kspir_context* c, int32_t max_batch, size_t workspace_budget_bytes, kspir_server** out) {
    return kspir_server_create_lanes(c, max_batch, workspace_budget_bytes, 1, out);
}

int kspir_server_submit(kspir_server* s, const kspir_database* db, const kspir_keys* keys, const uint64_t* query,
                        uint64_t* response, uint64_t* ticket) {
    if (!s || !query || !response || !ticket) return fail(KSPIR_EINVAL, "null argument");
    int rc = check_pair(s->ctx, db, keys);
    if (rc) return rc;
    const size_t one = batch_footprint_bytes(s->ctx, db->r, db->n1, 1);
    if (one > s->budget)
        return fail(KSPIR_EBUDGET, "one query of this database needs " + std::to_string(one) +
                                       " bytes of device workspace, above the lane budget of " +
                                       std::to_string(s->budget));
    {
        std::lock_guard<std::mutex> lock(s->mu);
        if (s->stop) return fail(KSPIR_EINVAL, "server is stopping");
        *ticket = s->next_id++;
        s->live.insert(*ticket);
        s->pending.push_back(kspir_server::Req{*ticket, db, keys, query, response, std::chrono::steady_clock::now()});
    }
    s->cv_work.notify_one();
    return KSPIR_OK;
}

int kspir_server_wait(kspir_server* s, uint64_t ticket, double* latency_ms) {
    if (!s) return fail(KSPIR_EINVAL, "null server");
    std::unique_lock<std::mutex> lock(s->mu);
    // each ticket is collected once: a second wait (or a ticket never
    // issued) is an error rather than a wait for a result that cannot come
    if (!s->live.count(ticket)) return fail(KSPIR_EINVAL, "unknown ticket, or ticket already collected");
    s->cv_done.wait(lock, [&] { return s->done.count(ticket) != 0; });
    const kspir_server::Done d = s->done[ticket];
    s->done.erase(ticket);
    s->live.erase(ticket);
    if (latency_ms) *latency_ms = d.ms;
    if (d.rc == KSPIR_ECANCELED) return fail(d.rc, "request was cancelled");
    if (d.rc == KSPIR_EINJECTED) return fail(d.rc, "batch failed on purpose (serve_fail_batches)");
    if (d.rc != KSPIR_OK) return fail(d.rc, "batch answer failed in the serving loop");
    return KSPIR_OK;
}

int kspir_server_cancel(kspir_server* s, uint64_t ticket) {
    if (!s) return fail(KSPIR_EINVAL, "null server");
    {
        std::lock_guard<std::mutex> lock(s->mu);
        if (!s->live.count(ticket)) return fail(KSPIR_EINVAL, "unknown ticket, or ticket already collected");
        auto it = std::find_if(s->pending.begin(), s->pending.end(),
                               [&](const kspir_server::Req& q) { return q.id == ticket; });
        if (it == s->pending.end()) return fail(KSPIR_EBUSY, "request already taken by a batch");
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - it->t0).count();
        s->pending.erase(it);
        s->done[ticket] = kspir_server::Done{KSPIR_ECANCELED, ms};
    }
    s->cv_done.notify_all();
    return KSPIR_OK;
}

int kspir_server_stats(kspir_server* s, uint64_t* batches, uint64_t* answered, uint64_t* pending) {
    if (!s) return fail(KSPIR_EINVAL, "null server");
    std::lock_guard<std::mutex> lock(s->mu);
    if (batches) *batches = s->batches;
    if (answered) *answered = s->answered;
    if (pending) *pending = s->pending.size();
    return KSPIR_OK;
}

int kspir_server_spanning_batches(kspir_server* s, uint64_t* spanning) {
    if (!s || !spanning) return fail(KSPIR_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    *spanning = s->spanning;
    return KSPIR_OK;
}

void kspir_server_destroy(kspir_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& L : s->lanes)
        if (L.worker.joinable()) L.worker.join();
    for (auto& L : s->lanes) {
        hipHostFree(L.qh);
        hipHostFree(L.rh);
        if (L.owned) kspir_context_destroy(L.ctx);
    }
    delete s;
}

int kspir_test_ntt(kspir_context* c, uint32_t* polys, int count, int prime, int inverse) {
    if (!c || !polys || count < 1 || prime < 0 || prime > 4) return fail(KSPIR_EINVAL, "bad argument");
    HIP_OK(hipSetDevice(c->device));
    uint32_t* d = nullptr;
    HIP_OK(hipMalloc(&d, (size_t)count * kN * 4));
    HIP_OK(hipMemcpy(d, polys, (size_t)count * kN * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_ntt_test, dim3(count), dim3(kNttThreads), 0, c->stream, d, prime, inverse, c->T);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(polys, d, (size_t)count * kN * 4, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail_hip("k_ntt_test", e);
    return KSPIR_OK;
}

}  // extern "C"
