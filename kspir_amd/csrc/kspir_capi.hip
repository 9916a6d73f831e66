// kspir_capi.hip — host side of the MI355X KsPIR library: device tables,
// resident database and key handles, and the launch sequence of the online
// answer behind the C ABI of include/kspir_amd.h.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kspir_kernels.h"
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

// ------------------------------------------------------------------ handles

struct kspir_context {
    int device = 0;
    hipStream_t stream = nullptr;
    Tabs T{};
    std::vector<void*> tables;
    // per-query workspace (grown on demand)
    int ws_r = 0, ws_n1 = 0, ws_n2 = 0;
    void* ws = nullptr;
    size_t ws_bytes = 0;
    uint64_t* q_dev = nullptr;      // [4][N]
    uint64_t* resp_dev = nullptr;   // [max(r,1)][2][N]
    uint32_t *D0, *D1, *bhat0, *mid, *icoef, *ctw, *dbout, *gcoef, *gks, *fd, *ecoef, *epart;
    uint32_t *extA, *extB, *pcoef, *ppart;
    uint64_t* u64tmp = nullptr;     // [max(r,1)][2][N]
};

struct kspir_database {
    kspir_context* ctx;
    int r, n1, n2;
    uint64_t* words;   // [r][N2][N1][N]
};

struct kspir_keys {
    kspir_context* ctx;
    int r, n1, n2;
    uint32_t* baby = nullptr;      // [N1/2][3][2][3][N] Montgomery
    uint32_t* giant = nullptr;     // [N2-1][2][2][2][N]
    uint32_t* packing = nullptr;   // [log2 r][4][2][2][N]
    uint32_t* rgsw = nullptr;      // [4][2][2][N]
    uint16_t* perms = nullptr;     // [N1][N]
};

namespace {

int build_tables(kspir_context* c) {
    const uint32_t qs[3] = {kQ1, kQ2, kBs};
    const u64 root_q = 3375402822066082ULL;   // crt.h:81 root_of_unity_crt
    const u64 roots[3] = {root_q % kQ1, root_q % kQ2, min_root(kBs)};
    for (int p = 0; p < 3; ++p) {
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
        uint32_t* d = nullptr;
        HIP_OK(hipMalloc(&d, sizeof(uint32_t) * (4 * kN + 2 * kN)));
        c->tables.push_back(d);
        HIP_OK(hipMemcpy(d + 0 * kN, fw.data(), 4 * kN, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d + 1 * kN, fwsh.data(), 4 * kN, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d + 2 * kN, iv.data(), 4 * kN, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d + 3 * kN, ivsh.data(), 4 * kN, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d + 4 * kN, pw.data(), 8 * kN, hipMemcpyHostToDevice));
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

// carve the per-query workspace; grows to the largest (r, N1, N2) seen
int ensure_workspace(kspir_context* c, int r, int n1) {
    const int n2 = kN / 2 / n1;
    if (c->ws && r <= c->ws_r && n1 <= c->ws_n1 && n2 <= c->ws_n2) return KSPIR_OK;
    const int R = r > c->ws_r ? r : c->ws_r;
    const int N1 = n1 > c->ws_n1 ? n1 : c->ws_n1;
    const int N2 = n2 > c->ws_n2 ? n2 : c->ws_n2;
    const size_t n = kN;
    const size_t rh = R / 2 > 0 ? R / 2 : 1;
    // sizes in u32 words
    const size_t sz_q = 2 * 4 * n, sz_resp = 2 * (size_t)(R > 2 ? R : 2) * 2 * n;   // u64 [max(R,2)][2][N]
    std::vector<std::pair<size_t, uint32_t**>> parts = {
        {9 * n, &c->D0}, {9 * n, &c->D1}, {3 * n, &c->bhat0}, {6 * n, &c->mid}, {3 * n, &c->icoef},
        {(size_t)N1 * 2 * 2 * n, &c->ctw},
        {(size_t)R * N2 * 4 * n, &c->dbout}, {(size_t)R * N2 * 2 * n, &c->gcoef},
        {(size_t)R * N2 * 4 * n, &c->gks}, {(size_t)R * 4 * n, &c->fd}, {(size_t)R * 4 * n, &c->ecoef},
        {(size_t)R * 16 * n, &c->epart}, {(size_t)R * 4 * n, &c->extA}, {(size_t)R * 4 * n, &c->extB},
        {rh * 2 * n, &c->pcoef}, {rh * 16 * n, &c->ppart},
    };
    size_t total = sz_q + 2 * sz_resp;
    for (auto& p : parts) total += (p.first + 63) / 64 * 64;
    if (c->ws) {
        HIP_OK(hipStreamSynchronize(c->stream));
        HIP_OK(hipFree(c->ws));
        c->ws = nullptr;
    }
    HIP_OK(hipMalloc(&c->ws, total * 4));
    c->ws_bytes = total * 4;
    uint32_t* base = (uint32_t*)c->ws;
    c->q_dev = (uint64_t*)base; base += sz_q;
    c->resp_dev = (uint64_t*)base; base += sz_resp;
    c->u64tmp = (uint64_t*)base; base += sz_resp;
    for (auto& p : parts) { *p.second = base; base += (p.first + 63) / 64 * 64; }
    c->ws_r = R;
    c->ws_n1 = N1;
    c->ws_n2 = N2;
    return KSPIR_OK;
}

int grid1d(size_t total) {
    size_t g = (total + 255) / 256;
    return (int)(g > 65535 ? 65535 : (g ? g : 1));
}

// The online answer for one query (test-pir.cpp:107-116) enqueued on `st`.
// q_dev: [4][N] u64 device; resp: [2][N] u64 device (may be null when only
// the first dimension is wanted).  Optional inspection outputs (device):
// fd_u64 [r][2][N] first dimension (NTT form), ext_u64 [r][2][N].
int enqueue_answer(kspir_context* c, const kspir_database* db, const kspir_keys* k, const uint64_t* q_dev,
                   uint64_t* resp, hipStream_t st, uint64_t* fd_u64, uint64_t* ext_u64, bool first_dim_only) {
    const int r = db->r, n1 = db->n1, n2 = db->n2, n1h = n1 / 2;
    const Tabs& T = c->T;
    // baby steps (twosteps.cpp:1206-1214, evalAutoSRNSSmallKeys)
    hipLaunchKernelGGL(k_baby_prep, dim3(16), dim3(kNttThreads), 0, st, q_dev, c->D0, c->bhat0, c->ctw, n1, T);
    LAUNCH_OK("k_baby_prep");
    if (n1h - 1 > 0) {
        hipLaunchKernelGGL(k_baby_rot, dim3(n1h - 1, 2), dim3(kNttThreads), 0, st, c->D0, c->bhat0, k->baby,
                           k->perms, c->ctw, c->mid, n1, 1, n1h - 2, T);
        LAUNCH_OK("k_baby_rot(first)");
    }
    hipLaunchKernelGGL(k_baby_mid_coef, dim3(3), dim3(kNttThreads), 0, st, c->mid, c->icoef, T);
    LAUNCH_OK("k_baby_mid_coef");
    hipLaunchKernelGGL(k_baby_mid_digits, dim3(9), dim3(kNttThreads), 0, st, c->icoef, c->D1, T);
    LAUNCH_OK("k_baby_mid_digits");
    hipLaunchKernelGGL(k_baby_rot, dim3(n1h, 2), dim3(kNttThreads), 0, st, c->D1, c->mid + 3 * kN, k->baby,
                       k->perms, c->ctw, (uint32_t*)nullptr, n1, n1h, -1, T);
    LAUNCH_OK("k_baby_rot(second)");
    // database pass (crt.cpp:643-768), one launch over all r blocks
    const size_t bw = (size_t)n2 * n1 * kN, obw = (size_t)n2 * 4 * kN;
    const dim3 dg(kN / 64, r);
    const uint64_t* ctw64 = (const uint64_t*)c->ctw;
    switch (n2 < 16 ? n2 : 16) {
        case 1: hipLaunchKernelGGL(k_db_product<1>, dg, dim3(64), 0, st, db->words, ctw64, c->dbout, n1, n2, bw, obw, T); break;
        case 2: hipLaunchKernelGGL(k_db_product<2>, dg, dim3(64), 0, st, db->words, ctw64, c->dbout, n1, n2, bw, obw, T); break;
        case 4: hipLaunchKernelGGL(k_db_product<4>, dg, dim3(64), 0, st, db->words, ctw64, c->dbout, n1, n2, bw, obw, T); break;
        case 8: hipLaunchKernelGGL(k_db_product<8>, dg, dim3(64), 0, st, db->words, ctw64, c->dbout, n1, n2, bw, obw, T); break;
        default: hipLaunchKernelGGL(k_db_product<16>, dg, dim3(64), 0, st, db->words, ctw64, c->dbout, n1, n2, bw, obw, T); break;
    }
    LAUNCH_OK("k_db_product");
    // giant steps (twosteps.cpp:599-677)
    if (n2 > 1) {
        hipLaunchKernelGGL(k_giant_coef, dim3(n2 - 1, r, 2), dim3(kNttThreads), 0, st, c->dbout, c->gcoef, n2, T);
        LAUNCH_OK("k_giant_coef");
        hipLaunchKernelGGL(k_giant_ks, dim3(n2 - 1, r, 2), dim3(kNttThreads), 0, st, c->dbout, c->gcoef, k->giant,
                           c->gks, n1, n2, T);
        LAUNCH_OK("k_giant_ks");
    }
    hipLaunchKernelGGL(k_ext_coef, dim3(r, 2, 2), dim3(kNttThreads), 0, st, c->dbout, c->gks, c->fd, c->ecoef, n2, T);
    LAUNCH_OK("k_ext_coef");
    if (fd_u64) {
        hipLaunchKernelGGL(k_rns_to_u64, dim3(grid1d((size_t)r * 2 * kN)), dim3(256), 0, st, c->fd, fd_u64, r, T);
        LAUNCH_OK("k_rns_to_u64");
    }
    if (first_dim_only) return KSPIR_OK;
    // external products (lwe.cpp:761-808)
    hipLaunchKernelGGL(k_ext_digit, dim3(r, 4, 2), dim3(kNttThreads), 0, st, c->ecoef, k->rgsw, c->epart, T);
    LAUNCH_OK("k_ext_digit");
    hipLaunchKernelGGL(k_ext_sum, dim3(grid1d((size_t)r * 4 * kN)), dim3(256), 0, st, c->epart, c->extA, r, T);
    LAUNCH_OK("k_ext_sum");
    if (ext_u64) {
        hipLaunchKernelGGL(k_response, dim3(2, r), dim3(kNttThreads), 0, st, c->extA, ext_u64, T);
        LAUNCH_OK("k_response(ext)");
    }
    // packing (answer.cpp:800-835)
    uint32_t* cur = c->extA;
    uint32_t* nxt = c->extB;
    for (int i = r / 2; i > 0; i >>= 1) {
        const int li = ilog2(i);
        const uint32_t g = (uint32_t)(kN / i + 1);
        const uint32_t ginv = (uint32_t)powmod(g, kN - 1, 2 * kN);   // odd g has order | N in (Z/2N)*
        const uint32_t* pkey = k->packing + (size_t)li * 4 * 2 * 2 * kN;
        hipLaunchKernelGGL(k_pack_coef, dim3(i, 2), dim3(kNttThreads), 0, st, cur, c->pcoef, i, T);
        LAUNCH_OK("k_pack_coef");
        hipLaunchKernelGGL(k_pack_digit, dim3(i, 4, 2), dim3(kNttThreads), 0, st, c->pcoef, pkey, c->ppart, ginv, T);
        LAUNCH_OK("k_pack_digit");
        hipLaunchKernelGGL(k_pack_combine, dim3(i, 2, 2), dim3(kNttThreads), 0, st, cur, c->ppart, nxt, i, g, T);
        LAUNCH_OK("k_pack_combine");
        uint32_t* t = cur; cur = nxt; nxt = t;
    }
    hipLaunchKernelGGL(k_response, dim3(2, 1), dim3(kNttThreads), 0, st, cur, resp, T);
    LAUNCH_OK("k_response");
    return KSPIR_OK;
}

int check_pair(const kspir_context* c, const kspir_database* db, const kspir_keys* k) {
    if (!c || !db || !k) return fail(KSPIR_EINVAL, "null handle");
    if (db->ctx != c || k->ctx != c) return fail(KSPIR_EINVAL, "handle belongs to another context");
    if (db->r != k->r || db->n1 != k->n1)
        return fail(KSPIR_EINVAL, "database and keys disagree on (r, N1)");
    if (!k->rgsw || (k->r > 1 && !k->packing)) return fail(KSPIR_EINVAL, "keys lack RGSW or packing keys");
    return KSPIR_OK;
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

const char* kspir_last_error(void) { return g_err.c_str(); }

int kspir_compute_permutation_matrix(int32_t* out, int32_t n1) {
    if (!out || !pow2(n1) || n1 > kN / 2) return fail(KSPIR_EINVAL, "n1 must be a power of two <= N/2");
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
    HIP_OK(hipSetDevice(device));
    kspir_context* c = new kspir_context();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail_hip("hipStreamCreate", e); }
    int rc = build_tables(c);
    if (rc != KSPIR_OK) { kspir_context_destroy(c); return rc; }
    *out = c;
    return KSPIR_OK;
}

void kspir_context_destroy(kspir_context* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (void* p : c->tables) hipFree(p);
    if (c->ws) hipFree(c->ws);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

void* kspir_context_stream(kspir_context* c) { return c ? (void*)c->stream : nullptr; }

int kspir_database_create(kspir_context* c, const uint64_t* datacrt, int32_t r, int32_t n1, int on_device,
                          kspir_database** out) {
    if (!c || !datacrt || !out) return fail(KSPIR_EINVAL, "null argument");
    if (!pow2(r) || r > kN || !pow2(n1) || n1 < 4 || n1 > kN / 2)
        return fail(KSPIR_EINVAL, "r must be a power of two <= N, N1 a power of two in [4, N/2]");
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

void kspir_database_destroy(kspir_database* db) {
    if (!db) return;
    hipSetDevice(db->ctx->device);
    hipStreamSynchronize(db->ctx->stream);
    hipFree(db->words);
    delete db;
}

int kspir_keys_create(kspir_context* c, int32_t r, int32_t n1, const uint64_t* baby, const uint64_t* giant,
                      const uint64_t* packing, const uint64_t* rgsw, const int32_t* permutations, kspir_keys** out) {
    if (!c || !baby || !permutations || !out) return fail(KSPIR_EINVAL, "null argument");
    if (!pow2(r) || r > kN || !pow2(n1) || n1 < 4 || n1 > kN / 2)
        return fail(KSPIR_EINVAL, "r must be a power of two <= N, N1 a power of two in [4, N/2]");
    const int n2 = kN / 2 / n1;
    if (n2 > 1 && !giant) return fail(KSPIR_EINVAL, "giant-step keys required when N2 > 1");
    if (r > 1 && !packing) return fail(KSPIR_EINVAL, "packing keys required when r > 1");
    std::vector<uint16_t> p16((size_t)n1 * kN);
    for (size_t i = 0; i < p16.size(); ++i) {
        if (permutations[i] < 0 || permutations[i] >= kN) return fail(KSPIR_EINVAL, "permutation entry out of range");
        p16[i] = (uint16_t)permutations[i];
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
    K_OK(hipMalloc(&k->baby, (size_t)nb * 3 * 2 * 3 * kN * 4));
    K_OK(hipMemcpy(raw, baby, (size_t)nb * 6 * 2 * kN * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_keys_baby, dim3(grid1d((size_t)nb * 6 * kN)), dim3(256), 0, c->stream, k->baby, raw, nb, c->T);
    K_OK(hipStreamSynchronize(c->stream));
    auto upload_q = [&](const uint64_t* src, int rows, uint32_t** dst) -> hipError_t {
        hipError_t ee = hipMalloc(dst, (size_t)rows * 2 * 2 * kN * 4);
        if (ee != hipSuccess) return ee;
        ee = hipMemcpy(raw, src, (size_t)rows * 2 * kN * 8, hipMemcpyHostToDevice);
        if (ee != hipSuccess) return ee;
        hipLaunchKernelGGL(k_keys_q, dim3(grid1d((size_t)rows * 2 * kN)), dim3(256), 0, c->stream, *dst, raw, rows, c->T);
        return hipStreamSynchronize(c->stream);
    };
    if (ng > 0) K_OK(upload_q(giant, ng * 2, &k->giant));
    if (np > 0) K_OK(upload_q(packing, np * 4, &k->packing));
    if (rgsw) K_OK(upload_q(rgsw, 4, &k->rgsw));
    K_OK(hipMalloc(&k->perms, p16.size() * 2));
    K_OK(hipMemcpy(k->perms, p16.data(), p16.size() * 2, hipMemcpyHostToDevice));
#undef K_OK
    *out = k;
    return cleanup(KSPIR_OK);
}

void kspir_keys_destroy(kspir_keys* k) {
    if (!k) return;
    hipSetDevice(k->ctx->device);
    hipStreamSynchronize(k->ctx->stream);
    hipFree(k->baby);
    hipFree(k->giant);
    hipFree(k->packing);
    hipFree(k->rgsw);
    hipFree(k->perms);
    delete k;
}

int kspir_answer_device(kspir_context* c, const kspir_database* db, const kspir_keys* k, const uint64_t* query_dev,
                        uint64_t* response_dev, void* stream) {
    int rc = check_pair(c, db, k);
    if (rc) return rc;
    if (!query_dev || !response_dev) return fail(KSPIR_EINVAL, "null buffer");
    HIP_OK(hipSetDevice(c->device));
    rc = ensure_workspace(c, db->r, db->n1);
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    return enqueue_answer(c, db, k, query_dev, response_dev, st, nullptr, nullptr, false);
}

int kspir_answer_inspect(kspir_context* c, const kspir_database* db, const kspir_keys* k, const uint64_t* query,
                         uint64_t* response, uint64_t* cipherbuf, uint64_t* firstdim, uint64_t* ext) {
    int rc = check_pair(c, db, k);
    if (rc) return rc;
    if (!query || !response) return fail(KSPIR_EINVAL, "null buffer");
    HIP_OK(hipSetDevice(c->device));
    rc = ensure_workspace(c, db->r, db->n1);
    if (rc) return rc;
    const int r = db->r;
    uint64_t* fd_dev = c->u64tmp;
    uint64_t* ext_buf = nullptr;
    if (ext) HIP_OK(hipMalloc(&ext_buf, (size_t)r * 2 * kN * 8));
    uint64_t* ext_dev = ext_buf;
    HIP_OK(hipMemcpyAsync(c->q_dev, query, 4 * kN * 8, hipMemcpyHostToDevice, c->stream));
    rc = enqueue_answer(c, db, k, c->q_dev, c->resp_dev, c->stream, firstdim ? fd_dev : nullptr,
                        ext ? ext_dev : nullptr, false);
    if (rc) { if (ext_buf) hipFree(ext_buf); return rc; }
    HIP_OK(hipMemcpyAsync(response, c->resp_dev, 2 * kN * 8, hipMemcpyDeviceToHost, c->stream));
    if (cipherbuf)
        HIP_OK(hipMemcpyAsync(cipherbuf, c->ctw, (size_t)kN * db->n1 * 2 * 8, hipMemcpyDeviceToHost, c->stream));
    if (firstdim) HIP_OK(hipMemcpyAsync(firstdim, fd_dev, (size_t)r * 2 * kN * 8, hipMemcpyDeviceToHost, c->stream));
    if (ext) HIP_OK(hipMemcpyAsync(ext, ext_dev, (size_t)r * 2 * kN * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (ext_buf) HIP_OK(hipFree(ext_buf));
    return KSPIR_OK;
}

int kspir_answer(kspir_context* c, const kspir_database* db, const kspir_keys* k, const uint64_t* query,
                 uint64_t* response) {
    return kspir_answer_inspect(c, db, k, query, response, nullptr, nullptr, nullptr);
}

int kspir_matrix_vector_mul_bsgs_rns_crt_large(kspir_context* c, const kspir_database* db, const kspir_keys* k,
                                               uint64_t* query, uint64_t* result) {
    if (!c || !db || !k) return fail(KSPIR_EINVAL, "null handle");
    if (db->ctx != c || k->ctx != c) return fail(KSPIR_EINVAL, "handle belongs to another context");
    if (db->r != k->r || db->n1 != k->n1) return fail(KSPIR_EINVAL, "database and keys disagree on (r, N1)");
    if (!query || !result) return fail(KSPIR_EINVAL, "null buffer");
    HIP_OK(hipSetDevice(c->device));
    int rc = ensure_workspace(c, db->r, db->n1);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(c->q_dev, query, 4 * kN * 8, hipMemcpyHostToDevice, c->stream));
    rc = enqueue_answer(c, db, k, c->q_dev, nullptr, c->stream, c->u64tmp, nullptr, true);
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
    int rc = ensure_workspace(c, k->r, k->n1);
    if (rc) return rc;
    hipStream_t st = c->stream;
    const Tabs& T = c->T;
    uint64_t* io = c->u64tmp;   // [0]: first_dim, [1]: result words
    HIP_OK(hipMemcpyAsync(io, first_dim, 2 * kN * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(io + 2 * kN, result, 2 * kN * 8, hipMemcpyHostToDevice, st));
    // first_dim in coefficient form -> ecoef planes [1][2][2][N]
    if (first_dim_is_ntt) {
        hipLaunchKernelGGL(k_u64_to_rns, dim3(1, 2, 2), dim3(kNttThreads), 0, st, io, c->fd, 0, T);
        LAUNCH_OK("k_u64_to_rns");
        hipLaunchKernelGGL(k_response, dim3(2, 1), dim3(kNttThreads), 0, st, c->fd, io, T);   // INTT -> u64
        LAUNCH_OK("k_response");
    }
    hipLaunchKernelGGL(k_u64_to_rns, dim3(1, 2, 2), dim3(kNttThreads), 0, st, io, c->ecoef, 0, T);
    LAUNCH_OK("k_u64_to_rns");
    hipLaunchKernelGGL(k_ext_digit, dim3(1, 4, 2), dim3(kNttThreads), 0, st, c->ecoef, k->rgsw, c->epart, T);
    LAUNCH_OK("k_ext_digit");
    hipLaunchKernelGGL(k_ext_sum, dim3(grid1d(4 * kN)), dim3(256), 0, st, c->epart, c->extA, 1, T);
    LAUNCH_OK("k_ext_sum");
    // result words are NTT-domain accumulators: add them
    hipLaunchKernelGGL(k_u64_to_rns, dim3(1, 2, 2), dim3(kNttThreads), 0, st, io + 2 * kN, c->extB, 0, T);
    LAUNCH_OK("k_u64_to_rns");
    hipLaunchKernelGGL(k_add_planes, dim3(grid1d(4 * kN)), dim3(256), 0, st, c->extA, c->extB, 1, T);
    LAUNCH_OK("k_add_planes");
    hipLaunchKernelGGL(k_response, dim3(2, 1), dim3(kNttThreads), 0, st, c->extA, io + 2 * kN, T);
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
    int rc = ensure_workspace(c, r, k->n1);
    if (rc) return rc;
    hipStream_t st = c->stream;
    const Tabs& T = c->T;
    HIP_OK(hipMemcpyAsync(c->u64tmp, rlwes, (size_t)r * 2 * kN * 8, hipMemcpyHostToDevice, st));
    // the packing tree works on NTT-domain planes
    hipLaunchKernelGGL(k_u64_to_rns, dim3(r, 2, 2), dim3(kNttThreads), 0, st, c->u64tmp, c->extA,
                       rlwes_are_ntt ? 0 : 1, T);
    LAUNCH_OK("k_u64_to_rns");
    uint32_t* cur = c->extA;
    uint32_t* nxt = c->extB;
    for (int i = r / 2; i > 0; i >>= 1) {
        const int li = ilog2(i);
        const uint32_t g = (uint32_t)(kN / i + 1);
        const uint32_t ginv = (uint32_t)powmod(g, kN - 1, 2 * kN);
        const uint32_t* pkey = k->packing + (size_t)li * 4 * 2 * 2 * kN;
        hipLaunchKernelGGL(k_pack_coef, dim3(i, 2), dim3(kNttThreads), 0, st, cur, c->pcoef, i, T);
        LAUNCH_OK("k_pack_coef");
        hipLaunchKernelGGL(k_pack_digit, dim3(i, 4, 2), dim3(kNttThreads), 0, st, c->pcoef, pkey, c->ppart, ginv, T);
        LAUNCH_OK("k_pack_digit");
        hipLaunchKernelGGL(k_pack_combine, dim3(i, 2, 2), dim3(kNttThreads), 0, st, cur, c->ppart, nxt, i, g, T);
        LAUNCH_OK("k_pack_combine");
        uint32_t* t = cur; cur = nxt; nxt = t;
    }
    hipLaunchKernelGGL(k_response, dim3(2, 1), dim3(kNttThreads), 0, st, cur, c->resp_dev, T);
    LAUNCH_OK("k_response");
    HIP_OK(hipMemcpyAsync(result, c->resp_dev, 2 * kN * 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return KSPIR_OK;
}

// ------------------------------------------------------------ runtime plumbing

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

int kspir_time_answer(kspir_context* c, const kspir_database* db, const kspir_keys* k, const uint64_t* query_dev,
                      uint64_t* response_dev, int warmup, int iters, float* ms) {
    if (!ms || iters < 1) return fail(KSPIR_EINVAL, "bad timing arguments");
    for (int i = 0; i < warmup; ++i) {
        int rc = kspir_answer_device(c, db, k, query_dev, response_dev, nullptr);
        if (rc) return rc;
    }
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; ++i) {
        int rc = kspir_answer_device(c, db, k, query_dev, response_dev, nullptr);
        if (rc) return rc;
    }
    HIP_OK(hipEventRecord(e1, c->stream));
    HIP_OK(hipEventSynchronize(e1));
    HIP_OK(hipEventElapsedTime(ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return KSPIR_OK;
}

int kspir_time_db_pass(kspir_context* c, const kspir_database* db, const kspir_keys* k, const uint64_t* query_dev,
                       int iters, float* ms_per_launch) {
    if (!ms_per_launch || iters < 1) return fail(KSPIR_EINVAL, "bad timing arguments");
    int rc = check_pair(c, db, k);
    if (rc) return rc;
    HIP_OK(hipSetDevice(c->device));
    rc = ensure_workspace(c, db->r, db->n1);
    if (rc) return rc;
    rc = enqueue_answer(c, db, k, query_dev, c->resp_dev, c->stream, nullptr, nullptr, true);
    if (rc) return rc;
    const int n1 = db->n1, n2 = db->n2, r = db->r;
    const size_t bw = (size_t)n2 * n1 * kN, obw = (size_t)n2 * 4 * kN;
    const dim3 dg(kN / 64, r);
    const uint64_t* ctw64 = (const uint64_t*)c->ctw;
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; ++i) {
        switch (n2 < 16 ? n2 : 16) {
            case 1: hipLaunchKernelGGL(k_db_product<1>, dg, dim3(64), 0, c->stream, db->words, ctw64, c->dbout, n1, n2, bw, obw, c->T); break;
            case 2: hipLaunchKernelGGL(k_db_product<2>, dg, dim3(64), 0, c->stream, db->words, ctw64, c->dbout, n1, n2, bw, obw, c->T); break;
            case 4: hipLaunchKernelGGL(k_db_product<4>, dg, dim3(64), 0, c->stream, db->words, ctw64, c->dbout, n1, n2, bw, obw, c->T); break;
            case 8: hipLaunchKernelGGL(k_db_product<8>, dg, dim3(64), 0, c->stream, db->words, ctw64, c->dbout, n1, n2, bw, obw, c->T); break;
            default: hipLaunchKernelGGL(k_db_product<16>, dg, dim3(64), 0, c->stream, db->words, ctw64, c->dbout, n1, n2, bw, obw, c->T); break;
        }
        LAUNCH_OK("k_db_product");
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

int kspir_test_ntt(kspir_context* c, uint32_t* polys, int count, int prime, int inverse) {
    if (!c || !polys || count < 1 || prime < 0 || prime > 2) return fail(KSPIR_EINVAL, "bad argument");
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
