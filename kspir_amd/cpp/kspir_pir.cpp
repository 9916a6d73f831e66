// kspir_pir.cpp — the drop-in C++ surface (kspir_pir.h) over the C ABI.
// Reference objects are flattened into the C ABI's layouts; databases and
// keys are uploaded on first use and cached per object, keyed by address and
// a hash of every word so that a changed or reallocated object is
// re-uploaded.
#include "kspir_pir.h"

#include <atomic>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <stdexcept>
#include <string>

#include "../../include/kspir_amd.h"
#include "../../include/kspir_amd_runtime.h"

using kspir_params::N;

RlweCiphertext::RlweCiphertext() : a(N, 0), b(N, 0) {}
RlweCiphertext::RlweCiphertext(uint64_t len, uint64_t module) : length(len), modulus(module), a(len, 0), b(len, 0) {}

RGSWCiphertext::RGSWCiphertext(uint64_t len, uint64_t module, uint64_t elln, uint64_t base, uint64_t bg)
    : length(len), modulus(module), ellnum(elln), PP(base), BBg(bg) {
    a.assign(2 * ellnum, std::vector<uint64_t>(len, 0));
    b.assign(2 * ellnum, std::vector<uint64_t>(len, 0));
}

AutoKey::AutoKey(int32_t len, uint64_t module, int32_t elln, uint64_t base, uint64_t bg)
    : length((uint64_t)len), modulus(module), ellnum(elln), PP(base), BBg(bg) {}

AutoKeyBSGSRNS::AutoKeyBSGSRNS(int32_t len, uint64_t module, uint64_t bsModule)
    : length((uint64_t)len), modulus(module), bsModulus(bsModule) {}

namespace {

std::mutex g_mu;
kspir_context* g_ctx = nullptr;

// While the library runs (device set-up included), rand() is pointed at a
// private state (initstate), and the caller's state is restored afterwards
// (setstate), so runtime code that draws rand() cannot move the caller's
// stream: the drop-ins advance it exactly as the reference does, and only
// by their own draws.
struct RandShield {
    char own[128];
    char* prev;
    RandShield() : prev(initstate(1, own, sizeof own)) {}
    ~RandShield() { setstate(prev); }
};

void check(int rc, const char* what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + kspir_last_error());
}

kspir_context* ctx() {
    if (!g_ctx) {
        const char* dev = std::getenv("KSPIR_DEVICE");
        check(kspir_context_create(dev ? std::atoi(dev) : 0, &g_ctx), "kspir_context_create");
    }
    return g_ctx;
}

bool pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

uint32_t pow5(int64_t e) {
    uint32_t r = 1;
    for (int64_t i = 0; i < e; ++i) r = (r * 5u) & (2 * N - 1);
    return r;
}

// 64-bit hash of n words, every word: four independent multiply-xorshift
// lanes (so the loop is not latency-bound), folded at the end
uint64_t hash_words(const uint64_t* p, size_t n, uint64_t seed) {
    uint64_t h[4] = {seed ^ 0x243F6A8885A308D3ULL, seed ^ 0x13198A2E03707344ULL, seed ^ 0xA4093822299F31D0ULL,
                     seed ^ 0x082EFA98EC4E6C89ULL};
    size_t i = 0;
    for (; i + 4 <= n; i += 4)
        for (int l = 0; l < 4; ++l) {
            uint64_t x = (h[l] ^ p[i + l]) * 0x9E3779B97F4A7C15ULL;
            h[l] = x ^ (x >> 29);
        }
    for (; i < n; ++i) {
        uint64_t x = (h[0] ^ p[i]) * 0x9E3779B97F4A7C15ULL;
        h[0] = x ^ (x >> 29);
    }
    uint64_t r = n;
    for (int l = 0; l < 4; ++l) {
        uint64_t x = (r ^ h[l]) * 0xBF58476D1CE4E5B9ULL;
        r = x ^ (x >> 31);
    }
    return r;
}

// the same over a large array, in chunks on the host's threads (a 1 GiB
// database hashes at memory speed)
uint64_t hash_words_parallel(const uint64_t* p, size_t n) {
    const size_t chunk = (size_t)1 << 22;   // 32 MiB of words per task
    const size_t nc = (n + chunk - 1) / chunk;
    if (nc <= 1) return hash_words(p, n, 1);
    std::vector<uint64_t> part(nc);
    unsigned nt = std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > 16) nt = 16;
    if (nt > nc) nt = (unsigned)nc;
    std::atomic<size_t> next(0);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (size_t c; (c = next.fetch_add(1)) < nc;) {
                const size_t b = c * chunk, e = b + chunk < n ? b + chunk : n;
                part[c] = hash_words(p + b, e - b, c + 1);
            }
        });
    for (auto& x : th) x.join();
    return hash_words(part.data(), nc, n);
}

struct Fp {
    uint64_t h = 1469598103934665603ULL;
    void add(uint64_t v) { h = (h ^ v) * 1099511628211ULL; }
    void poly(const std::vector<uint64_t>& p) {
        add(p.size());
        add(hash_words(p.data(), p.size(), 7));
    }
};

const std::vector<RlweCiphertext>& at(const std::map<int32_t, std::vector<RlweCiphertext> >& m, int32_t idx,
                                      size_t need) {
    const std::vector<RlweCiphertext>& v = m.at(idx);   // std::out_of_range as in the reference
    if (v.size() < need) throw std::runtime_error("key entry " + std::to_string(idx) + " has too few ciphertexts");
    return v;
}

void put_rlwe(uint64_t* dst, const RlweCiphertext& c) {
    if (c.a.size() != (size_t)N || c.b.size() != (size_t)N) throw std::runtime_error("RLWE of wrong length");
    std::memcpy(dst, c.a.data(), N * 8);
    std::memcpy(dst + N, c.b.data(), N * 8);
}

struct DbEntry { int32_t r, n1; uint64_t fp; kspir_database* h; bool pinned; };
struct KeyEntry { uint64_t fp; kspir_keys* h; };
std::map<const void*, DbEntry> g_dbs;
std::map<std::pair<const void*, uint64_t>, KeyEntry> g_keys;   // (object, shape tag)

std::map<const void*, bool> g_pinned;   // databases the caller declared unchanged

kspir_database* database(uint64_t* data, int32_t r, int32_t n1) {
    auto it = g_dbs.find(data);
    const bool pinned = g_pinned.count(data) != 0;
    if (it != g_dbs.end() && pinned && it->second.pinned && it->second.r == r && it->second.n1 == n1)
        return it->second.h;
    const uint64_t fp = hash_words_parallel(data, (size_t)r * N * (N / 2));
    if (it != g_dbs.end()) {
        if (it->second.r == r && it->second.n1 == n1 && it->second.fp == fp) {
            it->second.pinned = pinned;
            return it->second.h;
        }
        kspir_database_destroy(it->second.h);
        g_dbs.erase(it);
    }
    kspir_database* h = nullptr;
    check(kspir_database_create(ctx(), data, r, n1, 0, &h), "kspir_database_create");
    g_dbs[data] = DbEntry{r, n1, fp, h, pinned};
    return h;
}

kspir_keys* cached_keys(const void* obj, uint64_t tag, uint64_t fp, int32_t r, int32_t n1, const uint64_t* baby,
                        const uint64_t* giant, const uint64_t* packing, const uint64_t* rgsw, const int32_t* perms) {
    auto key = std::make_pair(obj, tag);
    auto it = g_keys.find(key);
    if (it != g_keys.end()) {
        if (it->second.fp == fp) return it->second.h;
        kspir_keys_destroy(it->second.h);
        g_keys.erase(it);
    }
    kspir_keys* h = nullptr;
    check(kspir_keys_create(ctx(), r, n1, baby, giant, packing, rgsw, perms, &h), "kspir_keys_create");
    g_keys[key] = KeyEntry{fp, h};
    return h;
}

// BSGS keys + permutation table of one AutoKeyBSGSRNS for (r, N1)
kspir_keys* bsgs_keys(const AutoKeyBSGSRNS& k, int32_t r, int32_t n1,
                      const std::vector<std::vector<int32_t> >& perms) {
    const int32_t n2 = N / 2 / n1;
    Fp f;
    f.add((uint64_t)r);
    f.add((uint64_t)n1);
    for (int32_t i = 1; i <= n1 / 2; ++i)
        for (const RlweCiphertext& c : at(k.keyMap, (int32_t)pow5(i), 6)) { f.poly(c.a); f.poly(c.b); }
    for (int32_t i = 1; i < n2; ++i)
        for (const RlweCiphertext& c : at(k.keyMap, (int32_t)pow5((int64_t)n1 * i), 2)) { f.poly(c.a); f.poly(c.b); }
    if ((int32_t)perms.size() < n1) throw std::runtime_error("permutations has fewer than N1 rows");
    for (int32_t i = 0; i < n1; ++i) {
        if (perms[i].size() != (size_t)N) throw std::runtime_error("permutation row of wrong length");
        for (size_t j = 0; j < perms[i].size(); ++j) f.add((uint64_t)(uint32_t)perms[i][j]);
    }
    const uint64_t tag = ((uint64_t)r << 32) | (uint32_t)n1;
    auto it = g_keys.find(std::make_pair((const void*)&k, tag));
    if (it != g_keys.end() && it->second.fp == f.h) return it->second.h;
    std::vector<uint64_t> baby((size_t)(n1 / 2) * 6 * 2 * N), giant((size_t)(n2 > 1 ? n2 - 1 : 0) * 2 * 2 * N);
    std::vector<int32_t> pm((size_t)n1 * N);
    for (int32_t i = 1; i <= n1 / 2; ++i) {
        const std::vector<RlweCiphertext>& v = at(k.keyMap, (int32_t)pow5(i), 6);
        for (int j = 0; j < 6; ++j) put_rlwe(&baby[(((size_t)(i - 1) * 6 + j) * 2) * N], v[j]);
    }
    for (int32_t i = 1; i < n2; ++i) {
        const std::vector<RlweCiphertext>& v = at(k.keyMap, (int32_t)pow5((int64_t)n1 * i), 2);
        for (int j = 0; j < 2; ++j) put_rlwe(&giant[(((size_t)(i - 1) * 2 + j) * 2) * N], v[j]);
    }
    for (int32_t i = 0; i < n1; ++i) std::memcpy(&pm[(size_t)i * N], perms[i].data(), N * 4);
    return cached_keys(&k, tag, f.h, r, n1, baby.data(), n2 > 1 ? giant.data() : nullptr, nullptr, nullptr, pm.data());
}

kspir_keys* rgsw_keys(const RGSWCiphertext& g) {
    if (g.getEllnum() != 2 || g.getBase() != (1ULL << 20) || g.getBg() != (1ULL << 18))
        throw std::runtime_error("externalProduct: RGSW gadget must be (ell 2, PP 2^20, Bg 2^18)");
    if (g.a.size() != 4 || g.b.size() != 4) throw std::runtime_error("RGSW must have 4 rows");
    Fp f;
    for (int i = 0; i < 4; ++i) { f.poly(g.a[i]); f.poly(g.b[i]); }
    auto it = g_keys.find(std::make_pair((const void*)&g, (uint64_t)0));
    if (it != g_keys.end() && it->second.fp == f.h) return it->second.h;
    std::vector<uint64_t> rg(4 * 2 * (size_t)N);
    for (int i = 0; i < 4; ++i) {
        if (g.a[i].size() != (size_t)N || g.b[i].size() != (size_t)N) throw std::runtime_error("RGSW row length");
        std::memcpy(&rg[(size_t)i * 2 * N], g.a[i].data(), N * 8);
        std::memcpy(&rg[(size_t)i * 2 * N + N], g.b[i].data(), N * 8);
    }
    return cached_keys(&g, 0, f.h, 1, 128, nullptr, nullptr, nullptr, rg.data(), nullptr);
}

kspir_keys* packing_keys(const AutoKey& k, int32_t r) {
    if (k.getEllnum() != 4 || k.getBase() != 0 || k.getBg() != (1ULL << 14))
        throw std::runtime_error("packingRLWEs: packing-key gadget must be (ell 4, PP 0, Bg 2^14)");
    Fp f;
    f.add((uint64_t)r);
    int lr = 0;
    for (int32_t i = 1; i < r; i <<= 1, ++lr)
        for (const RlweCiphertext& c : at(k.keyMap, N / i + 1, 4)) { f.poly(c.a); f.poly(c.b); }
    auto it = g_keys.find(std::make_pair((const void*)&k, (uint64_t)r));
    if (it != g_keys.end() && it->second.fp == f.h) return it->second.h;
    std::vector<uint64_t> pk((size_t)lr * 4 * 2 * N);
    lr = 0;
    for (int32_t i = 1; i < r; i <<= 1, ++lr) {
        const std::vector<RlweCiphertext>& v = at(k.keyMap, N / i + 1, 4);
        for (int j = 0; j < 4; ++j) put_rlwe(&pk[(((size_t)lr * 4 + j) * 2) * N], v[j]);
    }
    return cached_keys(&k, (uint64_t)r, f.h, r, 128, nullptr, nullptr, lr ? pk.data() : nullptr, nullptr, nullptr);
}

}  // namespace

void matrix_vector_mul_bsgs_rns_crt_large(std::vector<RlweCiphertext>& result, std::vector<RlweCiphertext>& input,
                                          uint64_t* data, const AutoKeyBSGSRNS& autoKey, const int32_t N1,
                                          std::vector<std::vector<int32_t> >& permutations, int32_t r) {
    std::lock_guard<std::mutex> lock(g_mu);
    RandShield shield;
    if (!data) throw std::runtime_error("matrix_vector_mul_bsgs_rns_crt_large: null database");
    if (!pow2(r) || !pow2(N1) || N1 < 32 || N1 > N / 2)
        throw std::runtime_error("matrix_vector_mul_bsgs_rns_crt_large: r and N1 must be powers of two, N1 in [32, N/2]");
    if (input.size() < 2) throw std::runtime_error("input needs the crtMod and bsMod parts");
    if (input[0].getIsNtt() || input[1].getIsNtt())
        throw std::runtime_error("input must be in coefficient form (query_bsgs_rns output)");
    if (result.size() < (size_t)r) throw std::runtime_error("result must hold r ciphertexts");
    kspir_database* db = database(data, r, N1);
    kspir_keys* keys = bsgs_keys(autoKey, r, N1, permutations);
    std::vector<uint64_t> q(4 * (size_t)N), res((size_t)r * 2 * N);
    put_rlwe(q.data(), input[0]);
    put_rlwe(q.data() + 2 * N, input[1]);
    check(kspir_matrix_vector_mul_bsgs_rns_crt_large(ctx(), db, keys, q.data(), res.data()),
          "matrix_vector_mul_bsgs_rns_crt_large");
    for (int32_t i = 0; i < r; ++i) {
        result[i].a.assign(res.begin() + (size_t)i * 2 * N, res.begin() + (size_t)i * 2 * N + N);
        result[i].b.assign(res.begin() + (size_t)i * 2 * N + N, res.begin() + (size_t)(i + 1) * 2 * N);
        result[i].setIsNtt(true);
    }
    // twosteps.cpp:1074-1086: input[0] becomes the modulus-switched query, NTT form
    input[0].a.assign(q.begin(), q.begin() + N);
    input[0].b.assign(q.begin() + N, q.begin() + 2 * N);
    input[0].setIsNtt(true);
}

void externalProduct(RlweCiphertext& result, RlweCiphertext& firstDim, const RGSWCiphertext& query) {
    std::lock_guard<std::mutex> lock(g_mu);
    RandShield shield;
    if (!query.getIsNtt()) throw std::runtime_error("externalProduct: RGSW must be in NTT form");   // lwe.cpp:763
    kspir_keys* keys = rgsw_keys(query);
    std::vector<uint64_t> res(2 * (size_t)N), fd(2 * (size_t)N);
    put_rlwe(res.data(), result);
    put_rlwe(fd.data(), firstDim);
    const bool was_ntt = firstDim.getIsNtt();
    check(kspir_external_product(ctx(), keys, res.data(), fd.data(), was_ntt ? 1 : 0), "externalProduct");
    result.a.assign(res.begin(), res.begin() + N);
    result.b.assign(res.begin() + N, res.end());
    result.setIsNtt(false);
    if (was_ntt) {
        firstDim.a.assign(fd.begin(), fd.begin() + N);
        firstDim.b.assign(fd.begin() + N, fd.end());
        firstDim.setIsNtt(false);
    }
}

void packingRLWEs(RlweCiphertext& result, std::vector<RlweCiphertext>& rlwes, const AutoKey& autokey) {
    std::lock_guard<std::mutex> lock(g_mu);
    RandShield shield;
    const int32_t r = (int32_t)rlwes.size();
    if (!pow2(r)) throw std::runtime_error("packingRLWEs: the number of RLWEs must be a power of two");
    const bool ntt = rlwes[0].getIsNtt();
    for (const RlweCiphertext& c : rlwes)
        if (c.getIsNtt() != ntt) throw std::runtime_error("packingRLWEs: RLWEs must all be in the same form");
    kspir_keys* keys = packing_keys(autokey, r);
    std::vector<uint64_t> rl((size_t)r * 2 * N), res(2 * (size_t)N);
    for (int32_t i = 0; i < r; ++i) put_rlwe(&rl[(size_t)i * 2 * N], rlwes[i]);
    check(kspir_packing_rlwes(ctx(), keys, res.data(), rl.data(), ntt ? 1 : 0), "packingRLWEs");
    result.a.assign(res.begin(), res.begin() + N);
    result.b.assign(res.begin() + N, res.end());
    result.setIsNtt(false);
}

void database_tobsgsntt(std::vector<std::vector<uint64_t> >& result, std::vector<std::vector<uint64_t> >& data,
                        uint64_t modulus, int32_t N1) {
    if (modulus != kspir_params::crtMod)
        throw std::runtime_error("database_tobsgsntt: the lift modulus must be crtMod (twosteps.cpp:155-166)");
    if (data.size() != (size_t)N || result.size() != (size_t)N / 2)
        throw std::runtime_error("database_tobsgsntt: data must be N x N/2, result N/2 x N");
    std::vector<uint64_t> flat((size_t)N * (N / 2)), out((size_t)(N / 2) * N);
    for (int i = 0; i < N; ++i) {
        if (data[i].size() < (size_t)N / 2) throw std::runtime_error("database_tobsgsntt: short data row");
        std::memcpy(&flat[(size_t)i * (N / 2)], data[i].data(), (N / 2) * 8);
    }
    std::lock_guard<std::mutex> lock(g_mu);
    RandShield shield;
    check(kspir_database_tobsgsntt(ctx(), flat.data(), N1, out.data()), "database_tobsgsntt");
    for (int i = 0; i < N / 2; ++i) result[i].assign(out.begin() + (size_t)i * N, out.begin() + (size_t)(i + 1) * N);
}

void database_tocrt(uint64_t* datacrt, std::vector<std::vector<uint64_t> >& data_ntt, int32_t N1) {
    const int32_t N2 = N / 2 / N1;
    for (int z = 0; z < N; ++z)
        for (int j = 0; j < N2; ++j)
            for (int k = 0; k < N1; ++k) {
                const uint64_t x = data_ntt[(size_t)j * N1 + k][z];
                datacrt[(size_t)z * N1 * N2 + (size_t)j * N1 + k] = (x % kspir_params::crtq1) | (x % kspir_params::crtq2) << 32;
            }
}

// ---------------------------------------------------------------- client

namespace {

// NTT of one polynomial on the GPU: mod crtMod as the CRT of the q1 and q2
// transforms (the HEXL root of crtMod is the CRT of their roots), mod bsMod
// directly
void ntt_host(std::vector<uint64_t>& v, uint64_t modulus, bool inverse) {
    using kspir_params::crtq1;
    using kspir_params::crtq2;
    std::vector<uint32_t> r((size_t)2 * N);
    if (modulus == kspir_params::crtMod) {
        for (int i = 0; i < N; ++i) {
            r[i] = (uint32_t)(v[i] % crtq1);
            r[N + i] = (uint32_t)(v[i] % crtq2);
        }
        check(kspir_test_ntt(ctx(), r.data(), 1, 0, inverse ? 1 : 0), "kspir_test_ntt");
        check(kspir_test_ntt(ctx(), r.data() + N, 1, 1, inverse ? 1 : 0), "kspir_test_ntt");
        // Garner: x = x1 + q1 * ((x2 - x1) * q1^-1 mod q2)
        uint64_t inv = 1, b = crtq1 % crtq2, e = crtq2 - 2;
        while (e) {
            if (e & 1) inv = (unsigned __int128)inv * b % crtq2;
            b = (unsigned __int128)b * b % crtq2;
            e >>= 1;
        }
        for (int i = 0; i < N; ++i) {
            const uint64_t x1 = r[i], x2 = r[N + i];
            const uint64_t t = (unsigned __int128)((x2 + crtq2 - x1 % crtq2) % crtq2) * inv % crtq2;
            v[i] = x1 + crtq1 * t;
        }
    } else if (modulus == kspir_params::bsMod) {
        for (int i = 0; i < N; ++i) r[i] = (uint32_t)v[i];
        check(kspir_test_ntt(ctx(), r.data(), 1, 2, inverse ? 1 : 0), "kspir_test_ntt");
        for (int i = 0; i < N; ++i) v[i] = r[i];
    } else {
        throw std::runtime_error("Secret: only crtMod and bsMod are supported");
    }
}

// rand() outputs in the reference's consumption order
std::vector<uint32_t> draw(size_t n) {
    std::vector<uint32_t> d(n);
    for (auto& x : d) x = (uint32_t)rand();
    return d;
}

const std::vector<uint64_t>& crt_secret(const Secret& s) {
    if (const_cast<Secret&>(s).getModulus() != kspir_params::crtMod)
        throw std::runtime_error("client keys need a secret mod crtMod");
    return s.coefficients();
}

RlweCiphertext rlwe_from(const uint64_t* p, uint64_t modulus, bool ntt) {
    RlweCiphertext c(N, modulus);
    c.a.assign(p, p + N);
    c.b.assign(p + N, p + 2 * N);
    c.setIsNtt(ntt);
    return c;
}

}  // namespace

Secret::Secret(uint64_t module, bool ntt) : modulus(module), coeff(N), nttform(false) {
    if (module != kspir_params::crtMod && module != kspir_params::bsMod)
        throw std::runtime_error("Secret: only crtMod and bsMod are supported");
    for (int i = 0; i < N; ++i) {   // secret.cpp:54-58
        const int32_t t = rand() % 3 - 1;
        coeff[i] = t < 0 ? (uint64_t)(t + (int64_t)module) : (uint64_t)t;
    }
    data = coeff;
    if (ntt) toNttForm();
}

void Secret::toNttForm() {
    if (nttform) return;
    std::lock_guard<std::mutex> g(g_mu);
    RandShield shield;
    data = coeff;
    ntt_host(data, modulus, false);
    nttform = true;
}

void Secret::toCoeffForm() {
    data = coeff;
    nttform = false;
}

void query_bsgs_rns(std::vector<RlweCiphertext>& cipher, Secret& queryKey, uint64_t row) {
    if (cipher.size() < 2) throw std::runtime_error("query_bsgs_rns: cipher needs the crtMod and bsMod ciphertexts");
    if (row >= (uint64_t)N / 2) throw std::runtime_error("query_bsgs_rns: row must be < N/2");
    const std::vector<uint64_t>& sc = crt_secret(queryKey);
    std::vector<uint32_t> d = draw((size_t)5 * N);   // sample_random mod Q, mod bsMod, sample_guass
    std::vector<uint64_t> out((size_t)4 * N);
    std::lock_guard<std::mutex> g(g_mu);
    RandShield shield;
    check(kspir_client_encrypt(ctx(), KSPIR_CLIENT_QUERY, sc.data(), (int64_t)row, 0, d.data(), d.size(), out.data()),
          "kspir_client_encrypt");
    cipher[0] = rlwe_from(out.data(), kspir_params::crtMod, false);
    cipher[1] = rlwe_from(out.data() + 2 * N, kspir_params::bsMod, false);
}

void RGSWCiphertext::keyGen(Secret& secret, const uint64_t index, bool is_reverse) {
    if (index >= length) throw std::runtime_error("RGSWCiphertext::keyGen: index must be < N");
    if (ellnum != 2 || PP != (1ULL << 20) || BBg != (1ULL << 18) || modulus != kspir_params::crtMod)
        throw std::runtime_error("RGSWCiphertext::keyGen: only (crtMod, ell 2, 2^20, 2^18) is supported");
    const std::vector<uint64_t>& sc = crt_secret(secret);
    std::vector<uint32_t> d = draw((size_t)12 * N);
    std::vector<uint64_t> out((size_t)8 * N);
    {
        std::lock_guard<std::mutex> g(g_mu);
    RandShield shield;
        check(kspir_client_encrypt(ctx(), is_reverse ? KSPIR_CLIENT_RGSW : KSPIR_CLIENT_RGSW_FORWARD, sc.data(),
                                   (int64_t)index, 0, d.data(), d.size(), out.data()),
              "kspir_client_encrypt");
    }
    for (int k = 0; k < 4; ++k) {
        a[k].assign(out.data() + (size_t)k * 2 * N, out.data() + (size_t)k * 2 * N + N);
        b[k].assign(out.data() + (size_t)k * 2 * N + N, out.data() + (size_t)(k + 1) * 2 * N);
    }
    isntt = true;
    secret.toNttForm();   // lwe.cpp:317-320
}

void AutoKeyBSGSRNS::bsgsKeyGen(Secret& secret, int32_t N1) {
    if (!pow2(N1) || N1 < 2 || N1 > N / 2) throw std::runtime_error("bsgsKeyGen: N1 must be a power of two <= N/2");
    const int32_t N2 = N / 2 / N1, nb = N1 / 2;
    const std::vector<uint64_t>& sc = crt_secret(secret);
    std::vector<uint32_t> d = draw((size_t)15 * N * nb + (size_t)6 * N * (N2 - 1));
    std::vector<uint64_t> out((size_t)nb * 12 * N + (size_t)(N2 - 1) * 4 * N);
    {
        std::lock_guard<std::mutex> g(g_mu);
    RandShield shield;
        check(kspir_client_encrypt(ctx(), KSPIR_CLIENT_BSGS, sc.data(), 0, N1, d.data(), d.size(), out.data()),
              "kspir_client_encrypt");
    }
    for (int i = 1; i <= nb; ++i) {
        std::vector<RlweCiphertext> v;
        for (int j = 0; j < 6; ++j)
            v.push_back(rlwe_from(out.data() + ((size_t)(i - 1) * 6 + j) * 2 * N,
                                  j < 3 ? kspir_params::crtMod : kspir_params::bsMod, true));
        keyMap.insert(std::make_pair((int32_t)pow5(i), v));
    }
    const uint64_t* g0 = out.data() + (size_t)nb * 12 * N;
    for (int i = 1; i < N2; ++i) {
        std::vector<RlweCiphertext> v;
        for (int j = 0; j < 2; ++j)
            v.push_back(rlwe_from(g0 + ((size_t)(i - 1) * 2 + j) * 2 * N, kspir_params::crtMod, true));
        keyMap.insert(std::make_pair((int32_t)pow5((int64_t)N1 * i), v));
    }
}

void AutoKey::keyGen(Secret& secret, const int32_t num, bool packing_rlwe) {
    if (num < 1 || num > N) throw std::runtime_error("AutoKey::keyGen: num must be in [1, N]");
    if (ellnum != 4 || PP != 0 || BBg != (1ULL << 14) || modulus != kspir_params::crtMod)
        throw std::runtime_error("AutoKey::keyGen: only (crtMod, ell 4, 0, 2^14) is supported");
    std::vector<int32_t> idx;
    if (packing_rlwe)
        for (int32_t i = 1; i < num; i <<= 1) idx.push_back(N / i + 1);
    else
        for (int32_t i = 2; i <= num + 1; i <<= 1) idx.push_back(i + 1);
    const std::vector<uint64_t>& sc = crt_secret(secret);
    std::vector<uint32_t> d = draw((size_t)12 * N * idx.size());
    std::vector<uint64_t> out((size_t)8 * N * idx.size());
    if (!idx.empty()) {
        std::lock_guard<std::mutex> g(g_mu);
    RandShield shield;
        check(kspir_client_encrypt(ctx(), packing_rlwe ? KSPIR_CLIENT_PACKING : KSPIR_CLIENT_PACKING_LWE, sc.data(),
                                   num, 0, d.data(), d.size(), out.data()),
              "kspir_client_encrypt");
    }
    for (size_t li = 0; li < idx.size(); ++li) {
        std::vector<RlweCiphertext> v;
        for (int j = 0; j < 4; ++j) v.push_back(rlwe_from(out.data() + (li * 4 + j) * 2 * N, modulus, true));
        keyMap.insert(std::make_pair(idx[li], v));
    }
}

void decrypt_bsgs_total(std::vector<uint64_t>& message, RlweCiphertext& cipher, Secret& secret,
                        const int32_t rlwesNum) {
    const std::vector<uint64_t>& sc = crt_secret(secret);
    if (cipher.a.size() != (size_t)N || cipher.b.size() != (size_t)N)
        throw std::runtime_error("decrypt_bsgs_total: RLWE of wrong length");
    std::vector<uint64_t> resp((size_t)2 * N);
    std::copy(cipher.a.begin(), cipher.a.end(), resp.begin());
    std::copy(cipher.b.begin(), cipher.b.end(), resp.begin() + N);
    message.assign(N, 0);
    std::lock_guard<std::mutex> g(g_mu);
    RandShield shield;
    if (cipher.getIsNtt()) {   // the C ABI takes coefficient form
        std::vector<uint64_t> a(cipher.a), b(cipher.b);
        ntt_host(a, kspir_params::crtMod, true);
        ntt_host(b, kspir_params::crtMod, true);
        std::copy(a.begin(), a.end(), resp.begin());
        std::copy(b.begin(), b.end(), resp.begin() + N);
    }
    check(kspir_decrypt_bsgs_total(ctx(), rlwesNum, 1, sc.data(), resp.data(), message.data()),
          "kspir_decrypt_bsgs_total");
    if (!cipher.getIsNtt()) {   // twosteps.cpp:393-401: the response is turned into NTT form
        ntt_host(cipher.a, kspir_params::crtMod, false);
        ntt_host(cipher.b, kspir_params::crtMod, false);
        cipher.setIsNtt(true);
    }
}

// ------------------------------------------------- database sampling, helpers

void sample_database_bsgs(std::vector<std::vector<uint64_t> >& data) {
    if (data.size() < (size_t)N) throw std::runtime_error("sample_database_bsgs: data needs N rows");
    for (int i = 0; i < N; ++i) {
        if (data[i].size() < (size_t)N / 2) throw std::runtime_error("sample_database_bsgs: short data row");
        for (int j = 0; j < N / 2; ++j) {
            uint64_t r;
            do {
                r = (uint64_t)rand() & 0xffff;
            } while (r > 65537ULL);   // twosteps.cpp:31-35 (bsgsp)
            data[i][j] = r;
        }
    }
}

void inverse_encode(std::vector<uint64_t>& message) {
    if (message.size() != (size_t)N) throw std::runtime_error("inverse_encode: message must hold N values");
    std::lock_guard<std::mutex> lock(g_mu);
    RandShield shield;
    check(kspir_inverse_encode(ctx(), message.data(), 1), "inverse_encode");
}

void compute_permutation_matrix(std::vector<std::vector<int32_t> >& permutations, const int32_t max_indexs,
                                const int32_t length) {
    if (length != N) throw std::runtime_error("compute_permutation_matrix: length must be N");
    if (max_indexs < 1 || max_indexs > N / 2)
        throw std::runtime_error("compute_permutation_matrix: max_indexs must be in [1, N/2]");
    if (permutations.size() < (size_t)max_indexs)
        throw std::runtime_error("compute_permutation_matrix: permutations has fewer than max_indexs rows");
    std::vector<int32_t> flat((size_t)max_indexs * N);
    check(kspir_compute_permutation_matrix(flat.data(), max_indexs), "compute_permutation_matrix");
    for (int32_t i = 0; i < max_indexs; ++i) permutations[i].assign(flat.begin() + (size_t)i * N, flat.begin() + (size_t)(i + 1) * N);
}

void showLargeVector(std::vector<uint64_t>& vals, std::string ss) {
    std::cout << ss;
    showLargeVector<uint64_t>(vals);
}

void kspir_pir_assume_unchanged(const void* p) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_pinned[p] = true;
}

void kspir_pir_forget(const void* p) {
    std::lock_guard<std::mutex> lock(g_mu);
    RandShield shield;
    g_pinned.erase(p);
    auto it = g_dbs.find(p);
    if (it != g_dbs.end()) {
        kspir_database_destroy(it->second.h);
        g_dbs.erase(it);
    }
    for (auto k = g_keys.begin(); k != g_keys.end();) {
        if (k->first.first == p) {
            kspir_keys_destroy(k->second.h);
            k = g_keys.erase(k);
        } else {
            ++k;
        }
    }
}
