// kspir_oracle.cpp — CPU oracle for the KsPIR online server answer.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++17, the reference
// algorithm of mmingluo/kspir for the online answer (BSGS first dimension,
// RGSW external product, RLWE packing), plus the client and offline steps
// that produce its inputs and the recover step that checks its output.  It is
// the word-level checker for the HIP implementation in kspir_amd/ and the
// reported CPU baseline.  Nothing in the product links against it.
//
// Each function cites the reference file:line it follows.  Arithmetic follows
// the reference's semantics literally, including the places where a value
// equal to the modulus (a non-canonical zero) can flow (automorphisms written
// as `q - x`, RlweCiphertext::multConst) and the `d > q` test in decompose.
// Intel HEXL (v1.2.5, not vendored) is replaced by its published semantics:
//   * NTT(N, q, root): negacyclic Cooley–Tukey over the table
//     w[bitrev(i)] = root^i, output slot i = a(root^(2*bitrev(i)+1));
//     ComputeInverse is the exact inverse; outputs fully reduced.
//   * NTT(N, q) uses MinimalPrimitiveRoot(2N, q): the smallest primitive
//     2N-th root of unity mod q.
//   * Where the CPU has AVX-512F/DQ the NTT runs 8 lanes at a time, as
//     HEXL's AVX-512 path does; the lane form gives the scalar form's words
//     (ko_ntt_set_vector switches between them for that check).
//   * EltwiseAddMod: s = x + y, s >= q ? s - q : s.
//     EltwiseSubMod: x >= y ? x - y : x + q - y.
//     EltwiseMultMod / EltwiseFMAMod: (x*y [+ z]) mod q, fully reduced.
// glibc rand() is called directly, in the reference's consumption order; the
// two draws of `((uint64_t)rand() << 31) | rand()` (samples.cpp:15) are taken
// left to right, which is the order clang (the reference's compiler,
// CMakeLists.txt:10) evaluates them in.

#include "kspir_oracle.h"

#include <immintrin.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using u64 = uint64_t;
using u128 = unsigned __int128;
using i64 = int64_t;
using i128 = __int128;

constexpr int N = 4096;                       // params.h:16
constexpr int LOGN = 12;
constexpr u64 Q1 = 268369921ULL;              // crt.h:16
constexpr u64 Q2 = 249561089ULL;              // crt.h:17
constexpr u64 Q = Q1 * Q2;                    // crt.h:18 crtMod
constexpr u64 BS = 16760833ULL;               // crt.h:32 bsMod
constexpr u64 PT = 65537ULL;                  // params.h:94 bsgsp
constexpr u64 DELTA = 1021937069741ULL;       // params.h:98 bsgsDelta
constexpr u64 ROOT_Q = 3375402822066082ULL;   // crt.h:81 root_of_unity_crt
constexpr u64 BS_INV_Q = 51309461009346113ULL;  // crt.h:33 bsModInv  (bsMod^-1 mod crtMod)
constexpr u64 Q_INV_BS = 3920321ULL;            // crt.h:34 crtModInv (crtMod^-1 mod bsMod)
// crt.cpp:25-26: b_inv_pa_i = (q2^-1 mod q1) * q2, pa_inv_b_i = (q1^-1 mod q2) * q1
constexpr u128 B_INV_PA = (u128)163640210ULL * Q2;
constexpr u128 PA_INV_B = (u128)97389680ULL * Q1;

int g_threads = 1;
thread_local double g_stage_us[5] = {0, 0, 0, 0, 0};   // per calling thread

// ---------------------------------------------------------------- arithmetic

// Barrett reduction of a 128-bit value by a modulus q < 2^62 using
// floor(2^128 / q); the quotient estimate is at most one short, so one
// conditional subtraction yields the exact residue (same construction as
// crt.cpp:345-366 barrett_raw_u128 / barrett_reduction_u128).
struct Mod {
    u64 q = 0, r_hi = 0, r_lo = 0;
    Mod() = default;
    explicit Mod(u64 q_) : q(q_) {
        u128 m = (~(u128)0) / q;   // floor((2^128 - 1) / q) == floor(2^128 / q) for odd q > 1
        r_hi = (u64)(m >> 64);
        r_lo = (u64)m;
    }
    u64 reduce(u128 x) const {
        u64 xl = (u64)x, xh = (u64)(x >> 64);
        u128 t0 = (u128)xl * r_lo;
        u128 t1 = (u128)xl * r_hi;
        u128 t2 = (u128)xh * r_lo;
        u128 mid = (t0 >> 64) + (u64)t1 + (u64)t2;
        u64 qhat = (u64)(t1 >> 64) + (u64)(t2 >> 64) + (u64)(mid >> 64) + xh * r_hi;
        u64 r = xl - qhat * q;
        return r >= q ? r - q : r;
    }
    u64 mul(u64 a, u64 b) const { return reduce((u128)a * b); }
    // HEXL EltwiseAddMod / EltwiseSubMod element semantics
    u64 add(u64 a, u64 b) const { u64 s = a + b; return s >= q ? s - q : s; }
    u64 sub(u64 a, u64 b) const { return a >= b ? a - b : a + q - b; }
};

u64 powmod(u64 b, u64 e, const Mod& m) {
    u64 r = 1 % m.q;
    b = b % m.q;
    while (e) {
        if (e & 1) r = m.mul(r, b);
        b = m.mul(b, b);
        e >>= 1;
    }
    return r;
}

u64 invmod(u64 a, u64 q) {
    i128 t = 0, nt = 1, r = q, nr = a % q;
    while (nr != 0) {
        i128 qq = r / nr;
        i128 tmp = t - qq * nt; t = nt; nt = tmp;
        tmp = r - qq * nr; r = nr; nr = tmp;
    }
    if (r != 1) throw std::runtime_error("not invertible");
    if (t < 0) t += q;
    return (u64)t;
}

inline u64 shoup_pre(u64 w, u64 q) { return (u64)(((u128)w << 64) / q); }
inline u64 mul_shoup(u64 x, u64 w, u64 wp, u64 q) {
    u64 qh = (u64)(((u128)x * wp) >> 64);
    u64 r = x * w - qh * q;
    return r >= q ? r - q : r;
}

inline uint32_t bitrev(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

// HEXL MinimalPrimitiveRoot(2N, q) semantics: smallest x with x^N == -1 mod q.
u64 min_root(u64 q) {
    Mod m(q);
    for (u64 x = 2; x < q; ++x)
        if (powmod(x, N, m) == q - 1) return x;
    throw std::runtime_error("no primitive root");
}

// 8 lanes of the scalar arithmetic above, for moduli below 2^62: the high
// 64 bits of a 64 x 64 product from four 32 x 32 products (AVX-512 has no
// 64-bit high multiply), Shoup's product as mul_shoup, and the corrections
// r >= q ? r - q : r written as an unsigned min of r and r - q (which wraps
// above r exactly when r < q).  Same words as the scalar code, lane by lane.
#define KO_AVX512 __attribute__((target("avx512f,avx512dq")))
KO_AVX512 inline __m512i v_mulhi64(__m512i a, __m512i b) {
    const __m512i lo32 = _mm512_set1_epi64(0xffffffffLL);
    const __m512i a1 = _mm512_srli_epi64(a, 32), b1 = _mm512_srli_epi64(b, 32);
    const __m512i p00 = _mm512_mul_epu32(a, b), p01 = _mm512_mul_epu32(a, b1);
    const __m512i p10 = _mm512_mul_epu32(a1, b), p11 = _mm512_mul_epu32(a1, b1);
    __m512i mid = _mm512_add_epi64(_mm512_srli_epi64(p00, 32), _mm512_and_si512(p01, lo32));
    mid = _mm512_add_epi64(mid, _mm512_and_si512(p10, lo32));
    __m512i hi = _mm512_add_epi64(p11, _mm512_srli_epi64(p01, 32));
    hi = _mm512_add_epi64(hi, _mm512_srli_epi64(p10, 32));
    return _mm512_add_epi64(hi, _mm512_srli_epi64(mid, 32));
}
KO_AVX512 inline __m512i v_mul_shoup(__m512i x, __m512i w, __m512i wp, __m512i q) {
    const __m512i qh = v_mulhi64(x, wp);
    const __m512i r = _mm512_sub_epi64(_mm512_mullo_epi64(x, w), _mm512_mullo_epi64(qh, q));
    return _mm512_min_epu64(r, _mm512_sub_epi64(r, q));
}
KO_AVX512 inline __m512i v_add(__m512i a, __m512i b, __m512i q) {
    const __m512i s = _mm512_add_epi64(a, b);
    return _mm512_min_epu64(s, _mm512_sub_epi64(s, q));
}
KO_AVX512 inline __m512i v_sub(__m512i a, __m512i b, __m512i q) {
    const __m512i d = _mm512_sub_epi64(a, b);
    return _mm512_min_epu64(d, _mm512_add_epi64(d, q));
}

// the NTT's lane form; off only to compare it with the scalar one
bool g_ntt_vector = true;
bool ntt_vector_usable() {
    static const bool ok = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq");
    return ok;
}

// Negacyclic NTT with HEXL conventions (see header comment).
struct Ntt {
    u64 q = 0, root = 0;
    Mod md;
    std::vector<u64> fw, fwp, iv, ivp;
    u64 ninv = 0, ninvp = 0;
    Ntt(u64 q_, u64 root_) : q(q_), root(root_), md(q_), fw(N), fwp(N), iv(N), ivp(N) {
        u64 p = 1;
        for (int i = 0; i < N; ++i) {
            uint32_t k = bitrev(i, LOGN);
            fw[k] = p;
            iv[k] = (i == 0) ? 1 : powmod(root, 2 * N - i, md);
            p = md.mul(p, root);
        }
        for (int i = 0; i < N; ++i) { fwp[i] = shoup_pre(fw[i], q); ivp[i] = shoup_pre(iv[i], q); }
        ninv = invmod(N, q);
        ninvp = shoup_pre(ninv, q);
    }
    void forward(u64* a) const {
        if (g_ntt_vector && ntt_vector_usable()) forward_v(a);
        else forward_scalar(a);
    }
    void inverse(u64* a) const {
        if (g_ntt_vector && ntt_vector_usable()) inverse_v(a);
        else inverse_scalar(a);
    }
    void forward_scalar(u64* a) const {
        for (int i = 0; i < N; ++i) a[i] %= q;   // a value equal to q enters as 0
        int t = N;
        for (int m = 1; m < N; m <<= 1) {
            t >>= 1;
            for (int i = 0; i < m; ++i) {
                const u64 w = fw[m + i], wp = fwp[m + i];
                u64* x = a + 2 * i * t;
                u64* y = x + t;
                for (int j = 0; j < t; ++j) {
                    u64 u = x[j], v = mul_shoup(y[j], w, wp, q);
                    x[j] = md.add(u, v);
                    y[j] = md.sub(u, v);
                }
            }
        }
    }
    void inverse_scalar(u64* a) const {
        for (int i = 0; i < N; ++i) a[i] %= q;
        int t = 1;
        for (int m = N / 2; m >= 1; m >>= 1) {
            for (int i = 0; i < m; ++i) {
                const u64 w = iv[m + i], wp = ivp[m + i];
                u64* x = a + 2 * i * t;
                u64* y = x + t;
                for (int j = 0; j < t; ++j) {
                    u64 u = x[j], v = y[j];
                    x[j] = md.add(u, v);
                    y[j] = mul_shoup(md.sub(u, v), w, wp, q);
                }
            }
            t <<= 1;
        }
        for (int i = 0; i < N; ++i) a[i] = mul_shoup(a[i], ninv, ninvp, q);
    }

    // The same butterflies 8 lanes at a time.  Stages with t >= 8 take x and
    // y as whole vectors; for t = 4, 2, 1 two vectors (16 words, 16 / 2t
    // groups) are split into their x and y halves by a two-source permute,
    // each lane given its group's twiddle, and put back by the inverse
    // permute.
    struct Lanes { __m512i x, y, back0, back1, tw; };
    KO_AVX512 static Lanes lanes(int t) {
        // x / y positions within 16 words, and the inverse permutes
        Lanes L;
        if (t == 4) {
            L.x = _mm512_setr_epi64(0, 1, 2, 3, 8, 9, 10, 11);
            L.y = _mm512_setr_epi64(4, 5, 6, 7, 12, 13, 14, 15);
            L.back0 = _mm512_setr_epi64(0, 1, 2, 3, 8, 9, 10, 11);
            L.back1 = _mm512_setr_epi64(4, 5, 6, 7, 12, 13, 14, 15);
            L.tw = _mm512_setr_epi64(0, 0, 0, 0, 1, 1, 1, 1);
        } else if (t == 2) {
            L.x = _mm512_setr_epi64(0, 1, 4, 5, 8, 9, 12, 13);
            L.y = _mm512_setr_epi64(2, 3, 6, 7, 10, 11, 14, 15);
            L.back0 = _mm512_setr_epi64(0, 1, 8, 9, 2, 3, 10, 11);
            L.back1 = _mm512_setr_epi64(4, 5, 12, 13, 6, 7, 14, 15);
            L.tw = _mm512_setr_epi64(0, 0, 1, 1, 2, 2, 3, 3);
        } else {
            L.x = _mm512_setr_epi64(0, 2, 4, 6, 8, 10, 12, 14);
            L.y = _mm512_setr_epi64(1, 3, 5, 7, 9, 11, 13, 15);
            L.back0 = _mm512_setr_epi64(0, 8, 1, 9, 2, 10, 3, 11);
            L.back1 = _mm512_setr_epi64(4, 12, 5, 13, 6, 14, 7, 15);
            L.tw = _mm512_setr_epi64(0, 1, 2, 3, 4, 5, 6, 7);
        }
        return L;
    }
    // words >= q reduced as the scalar pre-pass does
    KO_AVX512 void reduce_input(u64* a) const {
        const __m512i vq = _mm512_set1_epi64((long long)q);
        for (int i = 0; i < N; i += 8) {
            const __m512i v = _mm512_loadu_si512(a + i);
            if (_mm512_cmpge_epu64_mask(v, vq))
                for (int j = i; j < i + 8; ++j) a[j] %= q;
        }
    }
    KO_AVX512 void forward_v(u64* a) const {
        reduce_input(a);
        const __m512i vq = _mm512_set1_epi64((long long)q);
        int t = N;
        for (int m = 1; m < N; m <<= 1) {
            t >>= 1;
            if (t >= 8) {
                for (int i = 0; i < m; ++i) {
                    const __m512i w = _mm512_set1_epi64((long long)fw[m + i]);
                    const __m512i wp = _mm512_set1_epi64((long long)fwp[m + i]);
                    u64* x = a + 2 * i * t;
                    u64* y = x + t;
                    for (int j = 0; j < t; j += 8) {
                        const __m512i u = _mm512_loadu_si512(x + j);
                        const __m512i v = v_mul_shoup(_mm512_loadu_si512(y + j), w, wp, vq);
                        _mm512_storeu_si512(x + j, v_add(u, v, vq));
                        _mm512_storeu_si512(y + j, v_sub(u, v, vq));
                    }
                }
            } else {
                const Lanes L = lanes(t);
                const int per = 16 / (2 * t);   // groups per 16 words
                for (int b = 0; b < N; b += 16) {
                    const __m512i A = _mm512_loadu_si512(a + b), B = _mm512_loadu_si512(a + b + 8);
                    const int g0 = m + b / (2 * t);
                    const __m512i w = _mm512_permutexvar_epi64(L.tw, _mm512_maskz_loadu_epi64((1u << per) - 1, &fw[g0]));
                    const __m512i wp =
                        _mm512_permutexvar_epi64(L.tw, _mm512_maskz_loadu_epi64((1u << per) - 1, &fwp[g0]));
                    const __m512i u = _mm512_permutex2var_epi64(A, L.x, B);
                    const __m512i v = v_mul_shoup(_mm512_permutex2var_epi64(A, L.y, B), w, wp, vq);
                    const __m512i X = v_add(u, v, vq), Y = v_sub(u, v, vq);
                    _mm512_storeu_si512(a + b, _mm512_permutex2var_epi64(X, L.back0, Y));
                    _mm512_storeu_si512(a + b + 8, _mm512_permutex2var_epi64(X, L.back1, Y));
                }
            }
        }
    }
    KO_AVX512 void inverse_v(u64* a) const {
        reduce_input(a);
        const __m512i vq = _mm512_set1_epi64((long long)q);
        int t = 1;
        for (int m = N / 2; m >= 1; m >>= 1) {
            if (t >= 8) {
                for (int i = 0; i < m; ++i) {
                    const __m512i w = _mm512_set1_epi64((long long)iv[m + i]);
                    const __m512i wp = _mm512_set1_epi64((long long)ivp[m + i]);
                    u64* x = a + 2 * i * t;
                    u64* y = x + t;
                    for (int j = 0; j < t; j += 8) {
                        const __m512i u = _mm512_loadu_si512(x + j), v = _mm512_loadu_si512(y + j);
                        _mm512_storeu_si512(x + j, v_add(u, v, vq));
                        _mm512_storeu_si512(y + j, v_mul_shoup(v_sub(u, v, vq), w, wp, vq));
                    }
                }
            } else {
                const Lanes L = lanes(t);
                const int per = 16 / (2 * t);
                for (int b = 0; b < N; b += 16) {
                    const __m512i A = _mm512_loadu_si512(a + b), B = _mm512_loadu_si512(a + b + 8);
                    const int g0 = m + b / (2 * t);
                    const __m512i w = _mm512_permutexvar_epi64(L.tw, _mm512_maskz_loadu_epi64((1u << per) - 1, &iv[g0]));
                    const __m512i wp =
                        _mm512_permutexvar_epi64(L.tw, _mm512_maskz_loadu_epi64((1u << per) - 1, &ivp[g0]));
                    const __m512i u = _mm512_permutex2var_epi64(A, L.x, B), v = _mm512_permutex2var_epi64(A, L.y, B);
                    const __m512i X = v_add(u, v, vq), Y = v_mul_shoup(v_sub(u, v, vq), w, wp, vq);
                    _mm512_storeu_si512(a + b, _mm512_permutex2var_epi64(X, L.back0, Y));
                    _mm512_storeu_si512(a + b + 8, _mm512_permutex2var_epi64(X, L.back1, Y));
                }
            }
            t <<= 1;
        }
        const __m512i ni = _mm512_set1_epi64((long long)ninv), nip = _mm512_set1_epi64((long long)ninvp);
        for (int i = 0; i < N; i += 8)
            _mm512_storeu_si512(a + i, v_mul_shoup(_mm512_loadu_si512(a + i), ni, nip, vq));
    }
};

// secret.cpp:41-48, lwe.cpp:263-270: NTT(N, crtMod, root_of_unity_crt) for Q,
// NTT(N, q) (minimal root) for every other modulus.
const Ntt& ntt_q()  { static Ntt t(Q, ROOT_Q); return t; }
const Ntt& ntt_bs() { static Ntt t(BS, min_root(BS)); return t; }
const Ntt& ntt_p()  { static Ntt t(PT, min_root(PT)); return t; }
const Mod& mq()  { static Mod m(Q); return m; }
const Mod& mbs() { static Mod m(BS); return m; }
const Mod& mp()  { static Mod m(PT); return m; }

using Poly = std::vector<u64>;

struct Rlwe {
    Poly a, b;
    u64 q;
    bool isntt = false;
    explicit Rlwe(u64 q_ = Q) : a(N, 0), b(N, 0), q(q_) {}
};

// ------------------------------------------------------------------- samplers

// samples.cpp:9-18 sample_random
void sample_random(u64* a, u64 q, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        u64 hi = (u64)rand();
        u64 lo = (u64)rand();
        a[i] = ((hi << 31) | lo) % q;
    }
}
// samples.cpp:56-65 sample_guass (ternary)
void sample_guass(u64* r, u64 q) {
    for (int i = 0; i < N; ++i) {
        int32_t t = rand() % 3 - 1;
        r[i] = t < 0 ? q - 1 : (u64)t;
    }
}
// samples.cpp:67-76 guass_to_modulus
void guass_to_modulus(u64* r, u64 q1, u64 q2) {
    for (int i = 0; i < N; ++i)
        if (r[i] > q1 / 2) r[i] = (u64)((i64)r[i] - (i64)q1 + (i64)q2);
}

// ------------------------------------------------------------------ index maps

// utils.cpp:1446 pow_mod(int32): a^b mod a power of two (wrapping product)
int32_t pow_mod2(int32_t a, int32_t b, int32_t mod_number) {
    uint32_t r = 1;
    for (int32_t i = 0; i < b; ++i) r *= (uint32_t)a;
    return (int32_t)(r & (uint32_t)(mod_number - 1));
}

// utils.cpp:1473-1512 compute_hexl_rotate_indexes
void hexl_rotate_indexes(std::vector<int32_t>& h, std::vector<int32_t>& rot) {
    h.assign(N, 0);
    rot.assign(N, 0);
    h[0] = 1;
    for (int i = 0; i < LOGN; ++i) {
        int32_t c = 1 << i;
        int32_t interval = N / c;
        for (int32_t j = 0; j < c; ++j)
            h[j * interval + interval / 2] = h[j * interval] * (2 * c + 1) % (4 * c);
    }
    rot[0] = 1;
    rot[N - 1] = 2 * N - 1;
    for (int i = 1; i < N / 2; ++i) {
        rot[i] = rot[i - 1] * 5 % (2 * N);
        rot[N - 1 - i] = rot[N - i] * 5 % (2 * N);
    }
}

std::vector<int32_t> find_index_of(const std::vector<int32_t>& h) {
    std::vector<int32_t> f(N, 0);
    for (int i = 0; i < N; ++i) f[h[i] >> 1] = i;
    return f;
}

// utils.cpp:1530-1554 compute_inter_permutation
void inter_permutation(std::vector<int32_t>& inter, const std::vector<int32_t>& rot, int32_t index) {
    for (int32_t i = 0; i < N / 2 - index; ++i) {
        inter[rot[i] >> 1] = rot[i + index];
        inter[rot[N - 1 - i] >> 1] = rot[N - 1 - i - index];
    }
    for (int32_t i = N / 2 - index; i < N / 2; ++i) {
        inter[rot[i] >> 1] = rot[i + index - N / 2];
        inter[rot[N - 1 - i] >> 1] = rot[N - 1 - i - index + N / 2];
    }
}

// utils.cpp:1591-1619 compute_permutation_matrix, flattened [max_index][N]
std::vector<int32_t> permutation_matrix(int max_index) {
    std::vector<int32_t> h, rot;
    hexl_rotate_indexes(h, rot);
    std::vector<int32_t> f = find_index_of(h);
    std::vector<int32_t> inter(N, 0), out((size_t)max_index * N);
    for (int index = 0; index < max_index; ++index) {
        inter_permutation(inter, rot, index);
        for (int i = 0; i < N; ++i) out[(size_t)index * N + i] = f[inter[h[i] >> 1] >> 1];
    }
    return out;
}

// utils.cpp:1657-1675 compute_query_encode / 1677-1685 compute_query_decode
std::vector<int32_t> query_encode() {
    std::vector<int32_t> h, rot;
    hexl_rotate_indexes(h, rot);
    std::vector<int32_t> f = find_index_of(h), e(N);
    for (int i = 0; i < N; ++i) e[i] = f[rot[i] / 2];
    return e;
}
std::vector<int32_t> query_decode() {
    std::vector<int32_t> e = query_encode(), d(N);
    for (int i = 0; i < N; ++i) d[e[i]] = i;
    return d;
}

// utils.cpp:1355-1379 automorphic (canonical negation)
void automorphic(u64* out, const u64* in, int32_t index, u64 q) {
    for (int i = 0; i < N; ++i) {
        u64 dst = ((u64)i * (u64)index) & (2 * N - 1);
        if (dst >= (u64)N) out[dst - N] = (q - in[i]) % q;
        else out[dst] = in[i];
    }
}
// twosteps.cpp:757-783 automorphic_rns (maps q - 0 to 0)
void automorphic_rns(u64* out, const u64* in, int32_t index, u64 q) {
    for (int i = 0; i < N; ++i) {
        u64 dst = ((u64)i * (u64)index) & (2 * N - 1);
        if (dst >= (u64)N) { u64 t = q - in[i]; out[dst - N] = (t != q) ? t : 0; }
        else out[dst] = in[i];
    }
}
// answer.cpp:595-615 / twosteps.cpp:624-644: automorphism written as q - x
// (keeps q when x == 0)
void automorphic_raw(u64* out, const u64* in, int32_t index, u64 q) {
    for (int i = 0; i < N; ++i) {
        u64 dst = ((u64)i * (u64)index) & (2 * N - 1);
        if (dst >= (u64)N) out[dst - N] = q - in[i];
        else out[dst] = in[i];
    }
}

// --------------------------------------------------------- decompositions

// utils.cpp:346-359 powerOfBg
std::vector<u64> power_of_bg(u64 base, u64 bg, int ell) {
    std::vector<u64> r;
    if (base == 0) base = 1;
    r.push_back(base);
    for (int i = 0; i < ell - 1; ++i) r.push_back(base *= bg);
    return r;
}
// utils.cpp:361-375 powerOfBg128
std::vector<u128> power_of_bg128(u64 base, u64 bg, int ell) {
    std::vector<u128> r;
    if (base == 0) base = 1;
    u128 t = base;
    r.push_back(t);
    for (int i = 0; i < ell - 1; ++i) r.push_back(t *= bg);
    return r;
}

inline int ilog2(u64 x) { int b = 0; while (x > 1) { x >>= 1; ++b; } return b; }

// utils.cpp:754-811 decompose (unsigned digits with offset, `d > q` test)
void decompose(std::vector<Poly>& out, const u64* in, int ell, u64 base, u64 bg, u64 q) {
    const int gbits = ilog2(bg);
    const int basebits = base == 0 ? 0 : ilog2(base);
    const int nsub = 64 - gbits;
    u64 offset = 0;
    for (int i = 0; i < ell; ++i) offset += (u64)1 << (basebits + i * gbits);
    offset <<= (gbits - 1);
    offset += base / 2;
    for (int i = 0; i < N; ++i) {
        u64 d = in[i] + offset;
        if (d > q) d -= q;
        d >>= basebits;
        for (int j = 0; j < ell; ++j) {
            i64 r = (i64)(d << nsub >> nsub);
            r -= (i64)(bg / 2);
            d >>= gbits;
            out[j][i] = r >= 0 ? (u64)r : (u64)(r + (i64)q);
        }
    }
}

// utils.cpp:974-1044 decompose_rlwe (signed digits, arithmetic shifts)
void decompose_rlwe(std::vector<Poly>& out, const u64* in_a, const u64* in_b, int ell,
                    u64 base, u64 bg, u64 q) {
    const u64 half = q >> 1;
    const int gbits = ilog2(bg);
    const int basebits = ilog2(base);
    const int nsub = 64 - gbits;
    for (int part = 0; part < 2; ++part) {
        const u64* in = part == 0 ? in_a : in_b;
        for (int i = 0; i < N; ++i) {
            i64 d = in[i] > half ? (i64)(in[i] - q) : (i64)in[i];
            d >>= basebits;
            for (int j = 0; j < ell; ++j) {
                i64 r = (i64)((u64)d << nsub);
                r >>= nsub;
                d -= r;
                d >>= gbits;
                out[part * ell + j][i] = r >= 0 ? (u64)r : (u64)(r + (i64)q);
            }
        }
    }
}

// utils.cpp:516-532 crt_inv_bsgs with utils.cpp:406-432: the value in
// [0, Q*bsMod) congruent to x mod Q and y mod bsMod, computed as
// x*bsModInv*bsMod + y*crtModInv*Q mod Q*bsMod.
inline u128 crt_inv_bsgs(u64 x, u64 y) {
    const u128 M = (u128)Q * BS;
    u128 t1 = ((u128)x * BS_INV_Q) % M;
    t1 = (t1 * BS) % M;
    u128 t2 = ((u128)y * Q_INV_BS) % M;
    t2 = (t2 * Q) % M;
    return (t1 + t2) % M;
}

// utils.cpp:1120-1189 decompose_bsgs (80-bit unsigned digits with offset)
void decompose_bsgs(std::vector<Poly>& out_q, std::vector<Poly>& out_bs, const u64* in_q,
                    const u64* in_bs, int ell, u64 base, u64 bg) {
    const u128 M = (u128)Q * BS;
    const int gbits = ilog2(bg);
    const int basebits = base != 0 ? ilog2(base) : 0;
    const int nsub = 128 - gbits;
    u128 offset = 0;
    for (int i = 0; i < ell; ++i) offset += (u128)1 << (basebits + i * gbits);
    offset <<= (gbits - 1);
    offset += base / 2;
    for (int i = 0; i < N; ++i) {
        u128 d = crt_inv_bsgs(in_q[i], in_bs[i]) + offset;
        if (d > M) d -= M;
        d >>= basebits;
        for (int j = 0; j < ell; ++j) {
            i64 r = (i64)(u64)(d << nsub >> nsub);
            r -= (i64)(bg / 2);
            d >>= gbits;
            if (r >= 0) { out_q[j][i] = (u64)r; out_bs[j][i] = (u64)r; }
            else { out_q[j][i] = (u64)(r + (i64)Q); out_bs[j][i] = (u64)(r + (i64)BS); }
        }
    }
}

// crt.cpp:368-377 crt_compose: x*(q2^-1 mod q1)*q2 + y*(q1^-1 mod q2)*q1 mod Q
inline u64 crt_compose(u64 x, u64 y) {
    u128 v = (u128)x * B_INV_PA + (u128)y * PA_INV_B;
    return mq().reduce(v);
}

// twosteps.cpp:1367-1384 barrett_coeff == value mod q1 / mod q2
inline u64 pack_crt(u64 x) { return (x % Q1) | ((x % Q2) << 32); }

// ------------------------------------------------------------- element-wise

void ew_mul(u64* r, const u64* a, const u64* b, const Mod& m) { for (int i = 0; i < N; ++i) r[i] = m.mul(a[i], b[i]); }
void ew_add(u64* r, const u64* a, const u64* b, const Mod& m) { for (int i = 0; i < N; ++i) r[i] = m.add(a[i], b[i]); }
void ew_sub(u64* r, const u64* a, const u64* b, const Mod& m) { for (int i = 0; i < N; ++i) r[i] = m.sub(a[i], b[i]); }
void ew_mulc(u64* r, const u64* a, u64 c, const Mod& m) { for (int i = 0; i < N; ++i) r[i] = m.mul(a[i], c); }
// r -= a (.) b
void ew_submul(u64* r, const u64* a, const u64* b, const Mod& m) {
    for (int i = 0; i < N; ++i) r[i] = m.sub(r[i], m.mul(a[i], b[i]));
}

// --------------------------------------------------------------- encryption

struct Secret {
    Poly coeff;   // ternary, mod Q, coefficient form
    Poly ntt_q;   // NTT form mod Q (secret.cpp:56-59 / toNttForm)
};

// secret.cpp:13-60 Secret(crtMod, false)
Secret make_secret() {
    Secret s;
    s.coeff.resize(N);
    for (int i = 0; i < N; ++i) {
        int32_t t = rand() % 3 - 1;
        s.coeff[i] = t < 0 ? (u64)(t + (i64)Q) : (u64)t;
    }
    s.ntt_q = s.coeff;
    ntt_q().forward(s.ntt_q.data());
    return s;
}

// encrypt.cpp:53-104 encrypt(b, a, secret, message): a uniform in NTT form,
// b = a*s + NTT(e) + message (message in NTT form)
void encrypt_rlwe(u64* b, u64* a, const Secret& s, const u64* msg) {
    const Mod& m = mq();
    sample_random(a, Q, N);
    ew_mul(b, a, s.ntt_q.data(), m);
    Poly err(N);
    sample_guass(err.data(), Q);
    ntt_q().forward(err.data());
    ew_add(b, b, err.data(), m);
    ew_add(b, b, msg, m);
}

// encrypt.cpp:391-423 encrypt_special_rlwe (error multiplied by lwesnum)
void encrypt_special_rlwe(u64* b, u64* a, const Secret& s, const u64* msg, int32_t num) {
    const Mod& m = mq();
    sample_random(a, Q, N);
    ew_mul(b, a, s.ntt_q.data(), m);
    Poly err(N);
    sample_guass(err.data(), Q);
    ew_mulc(err.data(), err.data(), (u64)num, m);
    ntt_q().forward(err.data());
    ew_add(b, b, err.data(), m);
    ew_add(b, b, msg, m);
}

// secret mod bsMod in NTT form, and mod Q in NTT form (encrypt.cpp:471-477, 363-369)
void secret_rns_ntt(const Secret& s, Poly& s1, Poly& s2) {
    s1 = s.coeff;
    s2 = s.coeff;
    guass_to_modulus(s2.data(), Q, BS);
    ntt_q().forward(s1.data());
    ntt_bs().forward(s2.data());
}

// encrypt.cpp:443-495 encrypt_rns_bsgs with the overload at encrypt.cpp:118-152
void encrypt_rns_bsgs(Rlwe& c0, Rlwe& c1, const Secret& s, const Poly& message) {
    Poly s1, s2;
    secret_rns_ntt(s, s1, s2);
    const Mod &m1 = mq(), &m2 = mbs();
    sample_random(c0.a.data(), Q, N);
    sample_random(c1.a.data(), BS, N);
    ew_mul(c0.b.data(), c0.a.data(), s1.data(), m1);
    ew_mul(c1.b.data(), c1.a.data(), s2.data(), m2);
    Poly err(N), temp(N);
    sample_guass(err.data(), Q);
    for (int i = 0; i < N; ++i) temp[i] = m1.add(m1.mul(message[i], BS), err[i]);  // FMAMod(m, bsMod, err)
    ntt_q().forward(temp.data());
    ew_add(c0.b.data(), c0.b.data(), temp.data(), m1);
    guass_to_modulus(err.data(), Q, BS);
    temp = err;
    ntt_bs().forward(temp.data());
    ew_add(c1.b.data(), c1.b.data(), temp.data(), m2);
    ntt_q().inverse(c0.b.data());
    ntt_q().inverse(c0.a.data());
    ntt_bs().inverse(c1.b.data());
    ntt_bs().inverse(c1.a.data());
    c0.isntt = c1.isntt = false;
}

// encrypt.cpp:300-387 encrypt_rns_bsgs_autokey + encrypt_bsgs_autokey
void encrypt_rns_bsgs_autokey(u64* b1, u64* a1, u64* b2, u64* a2, const Secret& s,
                              const Poly& msg1, const Poly& msg2) {
    Poly s1, s2;
    secret_rns_ntt(s, s1, s2);
    const Mod &m1 = mq(), &m2 = mbs();
    sample_random(a1, Q, N);
    sample_random(a2, BS, N);
    ew_mul(b1, a1, s1.data(), m1);
    ew_mul(b2, a2, s2.data(), m2);
    Poly err(N), temp(N);
    sample_guass(err.data(), Q);
    ew_add(temp.data(), err.data(), msg1.data(), m1);
    ntt_q().forward(temp.data());
    ew_add(b1, b1, temp.data(), m1);
    guass_to_modulus(err.data(), Q, BS);
    ew_add(temp.data(), err.data(), msg2.data(), m2);
    ntt_bs().forward(temp.data());
    ew_add(b2, b2, temp.data(), m2);
}

// -------------------------------------------------------------- key material

// lwe.cpp:297-352 RGSWCiphertext(N, Q, 2, 2^20, 2^18).keyGen(secret, index, true)
void rgsw_keygen(u64* out /*[4][2][N]*/, const Secret& s, u64 index) {
    const int ell = 2;
    const Mod& m = mq();
    Poly msg(N, 0), temp(N);
    if (index == 0) msg[0] = 1;
    else msg[N - index] = Q - 1;
    std::vector<u64> mf = power_of_bg(1ULL << 20, 1ULL << 18, ell);
    for (int i = 0; i < ell; ++i) {
        ew_mulc(temp.data(), msg.data(), mf[i], m);
        ntt_q().forward(temp.data());
        u64* ra = out + (size_t)(ell + i) * 2 * N;
        encrypt_rlwe(ra + N, ra, s, temp.data());
        for (int k = 0; k < N; ++k) temp[k] = Q - temp[k];   // utils.cpp:294 negate
        ew_mul(temp.data(), temp.data(), s.ntt_q.data(), m);
        u64* rb = out + (size_t)i * 2 * N;
        encrypt_rlwe(rb + N, rb, s, temp.data());
    }
}

// lwe.cpp:1025-1070 generateSingleKeyBS (ell 3, PP 2^20, Bg 2^20)
void baby_key(u64* out /*[6][2][N]*/, int32_t index, const Secret& s) {
    const int ell = 3;
    Poly autosec(N), autosec2(N), t1(N), t2(N);
    std::vector<u128> mf = power_of_bg128(1ULL << 20, 1ULL << 20, ell);
    automorphic(autosec.data(), s.coeff.data(), index, Q);
    autosec2 = autosec;
    guass_to_modulus(autosec2.data(), Q, BS);
    for (int i = 0; i < ell; ++i) {
        ew_mulc(t1.data(), autosec.data(), (u64)(mf[i] % Q), mq());
        ew_mulc(t2.data(), autosec2.data(), (u64)(mf[i] % BS), mbs());
        u64* kq = out + (size_t)i * 2 * N;
        u64* kb = out + (size_t)(i + ell) * 2 * N;
        encrypt_rns_bsgs_autokey(kq + N, kq, kb + N, kb, s, t1, t2);
    }
}

// lwe.cpp:1072-1101 generateSingleKeyGS (ell 2, PP 2^20, Bg 2^18)
void giant_key(u64* out /*[2][2][N]*/, int32_t index, const Secret& s) {
    const int ell = 2;
    Poly autosec(N), temp(N);
    std::vector<u64> mf = power_of_bg(1ULL << 20, 1ULL << 18, ell);
    automorphic(autosec.data(), s.coeff.data(), index, Q);
    for (int i = 0; i < ell; ++i) {
        ew_mulc(temp.data(), autosec.data(), mf[i], mq());
        ntt_q().forward(temp.data());
        u64* k = out + (size_t)i * 2 * N;
        encrypt_rlwe(k + N, k, s, temp.data());
    }
}

// lwe.cpp:472-501 AutoKey::generateSingleKey(result, index, num, secret) (ell 4, PP 0, Bg 2^14)
void packing_key(u64* out /*[4][2][N]*/, int32_t index, int32_t num, const Secret& s) {
    const int ell = 4;
    Poly autosec(N), temp(N);
    std::vector<u64> mf = power_of_bg(0, 1ULL << 14, ell);
    automorphic(autosec.data(), s.coeff.data(), index, Q);
    for (int i = 0; i < ell; ++i) {
        ew_mulc(temp.data(), autosec.data(), mf[i], mq());
        ntt_q().forward(temp.data());
        u64* k = out + (size_t)i * 2 * N;
        encrypt_special_rlwe(k + N, k, s, temp.data(), num);
    }
}

// lwe.cpp:1143-1160 bsgsKeyGen
void bsgs_keygen(u64* baby, u64* giant, const Secret& s, int n1) {
    const int n2 = N / 2 / n1;
    for (int i = 1; i <= n1 / 2; ++i)
        baby_key(baby + (size_t)(i - 1) * 6 * 2 * N, pow_mod2(5, i, 2 * N), s);
    for (int i = 1; i < n2; ++i)
        giant_key(giant + (size_t)(i - 1) * 2 * 2 * N, pow_mod2(5, n1 * i, 2 * N), s);
}

// lwe.cpp:509-550 AutoKey::keyGen(secret, num, packing_rlwe = true)
void packing_keygen(u64* out, const Secret& s, int r) {
    int li = 0;
    for (int i = 1; i < r; i <<= 1, ++li)
        packing_key(out + (size_t)li * 4 * 2 * N, N / i + 1, r, s);
}

// twosteps.cpp:279-318 query_bsgs_rns
void query_bsgs_rns(u64* out /*[2][2][N]*/, const Secret& s, u64 row) {
    std::vector<int32_t> enc = query_encode();
    Poly msg(N, 0);
    msg[enc[row]] = 1;
    msg[enc[N - 1 - row]] = 1;
    ntt_p().inverse(msg.data());
    for (auto& me : msg) me = (u64)roundl((long double)me * Q / PT);
    Rlwe c0(Q), c1(BS);
    encrypt_rns_bsgs(c0, c1, s, msg);
    std::memcpy(out + 0 * N, c0.a.data(), N * 8);
    std::memcpy(out + 1 * N, c0.b.data(), N * 8);
    std::memcpy(out + 2 * N, c1.a.data(), N * 8);
    std::memcpy(out + 3 * N, c1.b.data(), N * 8);
}

// ------------------------------------------------------------ offline database

// twosteps.cpp:182-198 inverse_encode
void inverse_encode(Poly& msg) {
    ntt_p().forward(msg.data());
    std::vector<int32_t> dec = query_decode();
    Poly t = msg;
    for (int j = 0; j < N; ++j) msg[dec[j]] = t[j];
}

// twosteps.cpp:42-179 database_tobsgsntt; data: [N][N/2], result: [N/2][N]
void database_tobsgsntt(std::vector<u64>& result, const std::vector<u64>& data, int n1) {
    const int H = N / 2;
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < H; ++j) {
            int row = (j - i + H) & (H - 1);
            result[(size_t)i * N + j] = data[(size_t)row * H + j];
            result[(size_t)i * N + (N - 1 - j)] = data[(size_t)(N - 1 - row) * H + j];
        }
    std::vector<int32_t> enc = query_encode();
    std::vector<int32_t> perm = permutation_matrix(n1);
    const int n2 = H / n1;
    const u64 sub = Q - PT;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int kk = 0; kk < n2 * n1; ++kk) {
        int i = kk % n1;
        u64* row = &result[(size_t)kk * N];
        Poly t1(row, row + N), t2(N);
        for (int j = 0; j < N; ++j) t2[enc[j]] = t1[j];
        for (int j = 0; j < N; ++j) row[j] = t2[perm[(size_t)i * N + j]];
        ntt_p().inverse(row);
        for (int j = 0; j < N; ++j)
            if (row[j] > PT / 2) row[j] += sub;
        ntt_q().forward(row);
    }
}

// crt.cpp:859-885 database_tocrt
void database_tocrt(u64* datacrt, const std::vector<u64>& data_ntt, int n1) {
    const int n2 = N / 2 / n1;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < n2; ++j)
            for (int k = 0; k < n1; ++k)
                datacrt[(size_t)i * n1 * n2 + (size_t)j * n1 + k] = pack_crt(data_ntt[(size_t)(j * n1 + k) * N + i]);
}

// test-pir.cpp:46-63 one database block: sample, plant, encode column, preprocess
void make_db_block(u64* out, int n1, int k, int r, int plant, int tcol, int tpack,
                   std::vector<u64>& data, std::vector<u64>& data_ntt) {
    const int H = N / 2;
    // twosteps.cpp:23-39 sample_database_bsgs
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < H; ++j) {
            u64 v;
            do { v = (u64)(rand() & 0xffff); } while (v > PT);
            data[(size_t)i * H + j] = v;
        }
    if (plant) {
        for (int i = 0; i < N; i += r) data[(size_t)(i + tpack) * H + tcol] = (u64)(i + 1 + k);
        Poly rec(N);
        for (int i = 0; i < N; ++i) rec[i] = data[(size_t)i * H + tcol];
        inverse_encode(rec);
        for (int i = 0; i < N; ++i) data[(size_t)i * H + tcol] = rec[i];
    }
    database_tobsgsntt(data_ntt, data, n1);
    database_tocrt(out, data_ntt, n1);
}

// ------------------------------------------------------------ online answer

using Perms = const int32_t*;

// twosteps.cpp:937-1087 evalAutoSRNSFirst
void eval_auto_srns_first(std::vector<Rlwe>& result, Rlwe inter[2], Rlwe input[2], int n1h,
                          const u64* baby, Perms perms) {
    const int ell = 3;
    const u64 base = 1ULL << 20, bg = 1ULL << 20;
    const Mod &m1 = mq(), &m2 = mbs();
    std::vector<Rlwe> aux(n1h - 1, Rlwe(BS));
    std::vector<Poly> dec(ell, Poly(N)), dec_aux(ell, Poly(N));
    if (input[0].isntt) {
        ntt_q().inverse(input[0].a.data());
        ntt_bs().inverse(input[1].a.data());
        decompose_bsgs(dec, dec_aux, input[0].a.data(), input[1].a.data(), ell, base, bg);
        ntt_q().inverse(input[0].b.data());
        ntt_bs().inverse(input[1].b.data());
        input[0].isntt = input[1].isntt = false;
    } else {
        decompose_bsgs(dec, dec_aux, input[0].a.data(), input[1].a.data(), ell, base, bg);
    }
    for (int i = 0; i < ell; ++i) {
        ntt_q().forward(dec[i].data());
        ntt_bs().forward(dec_aux[i].data());
    }
#pragma omp parallel for num_threads(g_threads) schedule(dynamic)
    for (int i = 1; i < n1h; ++i) {
        // keyMap.at(5^i): baby key index i (1-based)
        const u64* key = baby + (size_t)(i - 1) * 6 * 2 * N;
        const int32_t* pm = perms + (size_t)i * N;
        Poly au1(N), au2(N);
        for (int j = 0; j < ell; ++j) {
            for (int k = 0; k < N; ++k) { au1[k] = dec[j][pm[k]]; au2[k] = dec_aux[j][pm[k]]; }
            const u64* kq = key + (size_t)j * 2 * N;
            const u64* kb = key + (size_t)(j + ell) * 2 * N;
            ew_submul(result[i - 1].a.data(), au1.data(), kq, m1);
            ew_submul(result[i - 1].b.data(), au1.data(), kq + N, m1);
            ew_submul(aux[i - 1].a.data(), au2.data(), kb, m2);
            ew_submul(aux[i - 1].b.data(), au2.data(), kb + N, m2);
        }
    }
#pragma omp parallel for num_threads(g_threads) schedule(dynamic)
    for (int i = 0; i < n1h - 1; ++i) {
        Poly t1(N), t2(N);
        ntt_q().inverse(result[i].a.data());
        ntt_q().inverse(result[i].b.data());
        ntt_bs().inverse(aux[i].a.data());
        ntt_bs().inverse(aux[i].b.data());
        int32_t idx = pow_mod2(5, i + 1, 2 * N);
        automorphic_rns(t1.data(), input[0].b.data(), idx, Q);
        automorphic_rns(t2.data(), input[1].b.data(), idx, BS);
        ew_add(result[i].b.data(), result[i].b.data(), t1.data(), m1);
        ew_add(aux[i].b.data(), aux[i].b.data(), t2.data(), m2);
        if (i == n1h - 2) {
            inter[0].a = result[i].a; inter[0].b = result[i].b; inter[0].isntt = result[i].isntt;
            inter[1].a = aux[i].a; inter[1].b = aux[i].b; inter[1].isntt = aux[i].isntt;
        }
        ew_sub(result[i].a.data(), result[i].a.data(), aux[i].a.data(), m1);
        ew_mulc(result[i].a.data(), result[i].a.data(), BS_INV_Q, m1);
        ntt_q().forward(result[i].a.data());
        ew_sub(result[i].b.data(), result[i].b.data(), aux[i].b.data(), m1);
        ew_mulc(result[i].b.data(), result[i].b.data(), BS_INV_Q, m1);
        ntt_q().forward(result[i].b.data());
        result[i].isntt = true;
    }
    ew_sub(input[0].a.data(), input[0].a.data(), input[1].a.data(), m1);
    ew_mulc(input[0].a.data(), input[0].a.data(), BS_INV_Q, m1);
    ntt_q().forward(input[0].a.data());
    ew_sub(input[0].b.data(), input[0].b.data(), input[1].b.data(), m1);
    ew_mulc(input[0].b.data(), input[0].b.data(), BS_INV_Q, m1);
    ntt_q().forward(input[0].b.data());
    input[0].isntt = true;
}

// twosteps.cpp:1090-1204 evalAutoSRNSSecond
void eval_auto_srns_second(std::vector<Rlwe>& result, Rlwe input[2], int n1h, const u64* baby,
                           Perms perms) {
    const int ell = 3;
    const u64 base = 1ULL << 20, bg = 1ULL << 20;
    const Mod &m1 = mq(), &m2 = mbs();
    std::vector<Rlwe> aux(n1h, Rlwe(BS));
    std::vector<Poly> dec(ell, Poly(N)), dec_aux(ell, Poly(N));
    if (input[0].isntt) {
        ntt_q().inverse(input[0].a.data());
        ntt_bs().inverse(input[1].a.data());
        decompose_bsgs(dec, dec_aux, input[0].a.data(), input[1].a.data(), ell, base, bg);
        ntt_q().inverse(input[0].b.data());
        ntt_bs().inverse(input[1].b.data());
        input[0].isntt = input[1].isntt = false;
    } else {
        decompose_bsgs(dec, dec_aux, input[0].a.data(), input[1].a.data(), ell, base, bg);
    }
    for (int i = 0; i < ell; ++i) {
        ntt_q().forward(dec[i].data());
        ntt_bs().forward(dec_aux[i].data());
    }
#pragma omp parallel for num_threads(g_threads) schedule(dynamic)
    for (int i = 0; i < n1h; ++i) {
        const u64* key = baby + (size_t)i * 6 * 2 * N;   // keyMap.at(5^(i+1))
        const int32_t* pm = perms + (size_t)(i + 1) * N;
        Poly au1(N), au2(N);
        for (int j = 0; j < ell; ++j) {
            for (int k = 0; k < N; ++k) { au1[k] = dec[j][pm[k]]; au2[k] = dec_aux[j][pm[k]]; }
            const u64* kq = key + (size_t)j * 2 * N;
            const u64* kb = key + (size_t)(j + ell) * 2 * N;
            ew_submul(result[n1h + i - 1].a.data(), au1.data(), kq, m1);
            ew_submul(result[n1h + i - 1].b.data(), au1.data(), kq + N, m1);
            ew_submul(aux[i].a.data(), au2.data(), kb, m2);
            ew_submul(aux[i].b.data(), au2.data(), kb + N, m2);
        }
    }
#pragma omp parallel for num_threads(g_threads) schedule(dynamic)
    for (int i = 0; i < n1h; ++i) {
        const int ri = n1h + i - 1;
        Poly t1(N), t2(N);
        ntt_q().inverse(result[ri].a.data());
        ntt_q().inverse(result[ri].b.data());
        ntt_bs().inverse(aux[i].a.data());
        ntt_bs().inverse(aux[i].b.data());
        int32_t idx = pow_mod2(5, i + 1, 2 * N);
        automorphic_rns(t1.data(), input[0].b.data(), idx, Q);
        automorphic_rns(t2.data(), input[1].b.data(), idx, BS);
        ew_add(result[ri].b.data(), result[ri].b.data(), t1.data(), m1);
        ew_add(aux[i].b.data(), aux[i].b.data(), t2.data(), m2);
        ew_sub(result[ri].a.data(), result[ri].a.data(), aux[i].a.data(), m1);
        ew_mulc(result[ri].a.data(), result[ri].a.data(), BS_INV_Q, m1);
        ntt_q().forward(result[ri].a.data());
        ew_sub(result[ri].b.data(), result[ri].b.data(), aux[i].b.data(), m1);
        ew_mulc(result[ri].b.data(), result[ri].b.data(), BS_INV_Q, m1);
        ntt_q().forward(result[ri].b.data());
        result[ri].isntt = true;
    }
}

// twosteps.cpp:1386-1403 reorientCipher
void reorient_cipher(u64* cbuf, const Rlwe& input, const std::vector<Rlwe>& rot, int n1) {
    for (int i = 0; i < N; ++i) {
        cbuf[(size_t)i * 2 * n1] = pack_crt(input.a[i]);
        cbuf[(size_t)i * 2 * n1 + 1] = pack_crt(input.b[i]);
        for (int j = 1; j < n1; ++j) {
            cbuf[(size_t)i * 2 * n1 + 2 * j] = pack_crt(rot[j - 1].a[i]);
            cbuf[(size_t)i * 2 * n1 + 2 * j + 1] = pack_crt(rot[j - 1].b[i]);
        }
    }
}

// crt.cpp:643-768 fastMultiplyQueryByDatabaseDim1InvCRT: u64 sums of the low
// (q1) and high (q2) 32-bit halves without per-term reduction, then
// crt_compose.  out: [n2][2N] (a in [0,N), b in [N,2N)).  Scalar form.
void db_product_scalar(u64* out, const u64* db, const u64* cbuf, int n1, int n2) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int z = 0; z < N; ++z) {
        const u64* ct = cbuf + (size_t)z * 2 * n1;
        for (int i = 0; i < n2; ++i) {
            const u64* d = db + (size_t)z * n2 * n1 + (size_t)i * n1;
            u64 sa0 = 0, sa1 = 0, sb0 = 0, sb1 = 0;
            for (int k = 0; k < n1; ++k) {
                u64 dl = d[k] & 0xffffffffULL, dh = d[k] >> 32;
                u64 ca = ct[2 * k], cb = ct[2 * k + 1];
                sa0 += dl * (ca & 0xffffffffULL);
                sa1 += dh * (ca >> 32);
                sb0 += dl * (cb & 0xffffffffULL);
                sb1 += dh * (cb >> 32);
            }
            out[(size_t)i * 2 * N + z] = crt_compose(sa0, sa1);
            out[(size_t)i * 2 * N + N + z] = crt_compose(sb0, sb1);
        }
    }
}

// The same sums in the shape of the reference's AVX-512 branch
// (crt.cpp:660-768): one 512-bit load holds the packed a|b words of four
// rotations, each database word is broadcast to its a and b lanes, and
// _mm512_mul_epu32 forms the q1 (low halves) and q2 (high halves) products,
// accumulated per lane without reduction.  Lanes 0, 2, 4, 6 carry a and 1, 3,
// 5, 7 carry b.  Exact for N1 <= 256 like the scalar form (every product is
// below 2^56).  Requires N1 % 4 == 0.
__attribute__((target("avx512f"))) void db_product_avx512(u64* out, const u64* db, const u64* cbuf, int n1, int n2) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int z = 0; z < N; ++z) {
        const u64* ct = cbuf + (size_t)z * 2 * n1;
        for (int i = 0; i < n2; ++i) {
            const u64* d = db + (size_t)z * n2 * n1 + (size_t)i * n1;
            __m512i s_lo = _mm512_setzero_si512(), s_hi = _mm512_setzero_si512();
            for (int k = 0; k < n1; k += 4) {
                const __m512i c = _mm512_loadu_si512(ct + 2 * k);
                const __m512i w = _mm512_set_epi64((long long)d[k + 3], (long long)d[k + 3], (long long)d[k + 2],
                                                   (long long)d[k + 2], (long long)d[k + 1], (long long)d[k + 1],
                                                   (long long)d[k], (long long)d[k]);
                s_lo = _mm512_add_epi64(s_lo, _mm512_mul_epu32(c, w));
                s_hi = _mm512_add_epi64(s_hi, _mm512_mul_epu32(_mm512_srli_epi64(c, 32), _mm512_srli_epi64(w, 32)));
            }
            alignas(64) u64 lo[8], hi[8];
            _mm512_store_si512(lo, s_lo);
            _mm512_store_si512(hi, s_hi);
            out[(size_t)i * 2 * N + z] = crt_compose(lo[0] + lo[2] + lo[4] + lo[6], hi[0] + hi[2] + hi[4] + hi[6]);
            out[(size_t)i * 2 * N + N + z] = crt_compose(lo[1] + lo[3] + lo[5] + lo[7], hi[1] + hi[3] + hi[5] + hi[7]);
        }
    }
}

// the AVX-512 form where the host has it, the scalar form otherwise
void db_product(u64* out, const u64* db, const u64* cbuf, int n1, int n2) {
    static const bool avx512 = __builtin_cpu_supports("avx512f");
    if (avx512 && n1 % 4 == 0) db_product_avx512(out, db, cbuf, n1, n2);
    else db_product_scalar(out, db, cbuf, n1, n2);
}

// twosteps.cpp:599-677 evalAutoRNSCRT (giant steps; ell 2, PP 2^20, Bg 2^18)
void eval_auto_rns_crt(Rlwe& result, u64* out /*[n2][2N]*/, int n2, const u64* giant) {
    const int ell = 2;
    const u64 base = 1ULL << 20, bg = 1ULL << 18;
    const int n1 = N / 2 / n2;
    const Mod& m = mq();
    Poly ta(N), tb(N), temp(N);
    std::vector<Poly> dec(ell, Poly(N));
    for (int c = 1; c < n2; ++c) {
        u64* ia = out + (size_t)c * 2 * N;
        u64* ib = ia + N;
        ntt_q().inverse(ia);
        ntt_q().inverse(ib);
        int32_t index = pow_mod2(5, n1 * c, 2 * N);
        automorphic_raw(ta.data(), ia, index, Q);
        automorphic_raw(tb.data(), ib, index, Q);
        decompose(dec, ta.data(), ell, base, bg, Q);
        const u64* key = giant + (size_t)(c - 1) * 2 * 2 * N;
        ntt_q().forward(tb.data());
        ew_add(result.b.data(), result.b.data(), tb.data(), m);
        for (int i = 0; i < ell; ++i) {
            ntt_q().forward(dec[i].data());
            const u64* k = key + (size_t)i * 2 * N;
            ew_submul(result.a.data(), dec[i].data(), k, m);
            ew_submul(result.b.data(), dec[i].data(), k + N, m);
        }
    }
}

// lwe.cpp:761-808 externalProduct (RGSW ell 2, PP 2^20, Bg 2^18); result accumulates
void external_product(Rlwe& result, Rlwe& fd, const u64* rgsw) {
    const int ell = 2;
    const Mod& m = mq();
    if (fd.isntt) {
        ntt_q().inverse(fd.a.data());
        ntt_q().inverse(fd.b.data());
        fd.isntt = false;
    }
    std::vector<Poly> dec(2 * ell, Poly(N));
    decompose_rlwe(dec, fd.a.data(), fd.b.data(), ell, 1ULL << 20, 1ULL << 18, Q);
    Poly temp(N);
    for (int i = 0; i < 2 * ell; ++i) {
        ntt_q().forward(dec[i].data());
        ew_mul(temp.data(), dec[i].data(), rgsw + (size_t)i * 2 * N, m);
        ew_add(result.a.data(), result.a.data(), temp.data(), m);
    }
    for (int i = 0; i < 2 * ell; ++i) {
        ew_mul(temp.data(), dec[i].data(), rgsw + (size_t)i * 2 * N + N, m);
        ew_add(result.b.data(), result.b.data(), temp.data(), m);
    }
    ntt_q().inverse(result.a.data());
    ntt_q().inverse(result.b.data());
    result.isntt = false;
}

// lwe.cpp:102-119 RlweCiphertext::multConst (negacyclic shift by X^number)
void mult_const(Rlwe& c, int32_t number) {
    Poly ta = c.a, tb = c.b;
    for (int i = 0; i < N - number; ++i) { c.a[i + number] = ta[i]; c.b[i + number] = tb[i]; }
    for (int i = N - number; i < N; ++i) {
        int idx = i + number - N;
        c.a[idx] = c.q - ta[i];
        c.b[idx] = c.q - tb[i];
    }
}

// answer.cpp:576-648 evalAuto(result, index, AutoKey) (ell 4, PP 0, Bg 2^14)
void eval_auto_packing(Rlwe& res, int32_t index, const u64* key /*[4][2][N]*/) {
    const int ell = 4;
    const Mod& m = mq();
    if (res.isntt) {
        ntt_q().inverse(res.a.data());
        ntt_q().inverse(res.b.data());
        res.isntt = false;
    }
    Poly ta(N), tb(N);
    for (int i = 0; i < N; ++i) {
        u64 dst = ((u64)i * (u64)index) & (2 * N - 1);
        if (dst >= (u64)N) { ta[dst - N] = Q - res.a[i]; tb[dst - N] = Q - res.b[i]; }
        else { ta[dst] = res.a[i]; tb[dst] = res.b[i]; }
        res.a[i] = 0;
    }
    std::vector<Poly> dec(ell, Poly(N));
    decompose(dec, ta.data(), ell, 0, 1ULL << 14, Q);
    res.b = tb;
    ntt_q().forward(res.b.data());
    for (int i = 0; i < ell; ++i) {
        ntt_q().forward(dec[i].data());
        const u64* k = key + (size_t)i * 2 * N;
        ew_submul(res.a.data(), dec[i].data(), k, m);
        ew_submul(res.b.data(), dec[i].data(), k + N, m);
    }
    ntt_q().inverse(res.a.data());
    ntt_q().inverse(res.b.data());
}

// answer.cpp:720-739 convertTwoRLWEs
void convert_two_rlwes(Rlwe& result, Rlwe& cipher, u64 twol, u64 nums, const u64* packing) {
    const Mod& m = mq();
    int32_t x = (int32_t)(nums / twol);
    mult_const(cipher, x);
    Rlwe temp = result;
    ew_sub(result.a.data(), result.a.data(), cipher.a.data(), m);
    ew_sub(result.b.data(), result.b.data(), cipher.b.data(), m);
    ew_add(cipher.a.data(), cipher.a.data(), temp.a.data(), m);
    ew_add(cipher.b.data(), cipher.b.data(), temp.b.data(), m);
    int32_t index = N / x + 1;
    // packing key for index N/i + 1 sits at level log2(i)
    int li = ilog2((u64)x);
    eval_auto_packing(result, index, packing + (size_t)li * 4 * 2 * N);
    ew_add(result.a.data(), result.a.data(), cipher.a.data(), m);
    ew_add(result.b.data(), result.b.data(), cipher.b.data(), m);
}

// answer.cpp:800-835 packingRLWEs
void packing_rlwes(Rlwe& result, std::vector<Rlwe>& rlwes, const u64* packing) {
    const u64 num = rlwes.size();
    for (auto& c : rlwes)
        if (c.isntt) {
            ntt_q().inverse(c.a.data());
            ntt_q().inverse(c.b.data());
            c.isntt = false;
        }
    for (u64 i = num / 2; i > 0; i >>= 1)
        for (u64 j = 0; j < i; ++j)
            convert_two_rlwes(rlwes[j], rlwes[i + j], num / i, num, packing);
    result = rlwes[0];
}

// ------------------------------------------- offline giant-key derivation

constexpr u64 AUX = 268361729ULL;                 // crt.h:37 auxMod
constexpr u64 AUX_INV_Q = 48735363086513670ULL;   // crt.h:56 auxModInv (auxMod^-1 mod crtMod)
constexpr u64 Q_INV_AUX = 73083388ULL;            // crt.h:63 crtModInvAux (crtMod^-1 mod auxMod)

// lwe.cpp:1016-1019: NTT(length, auxMod) takes the minimal root
const Ntt& ntt_aux() { static Ntt t(AUX, min_root(AUX)); return t; }
const Mod& maux() { static Mod m(AUX); return m; }

// utils.cpp:462-488 compute_mult_qinvqi_bsgs_aux: input * qinv * qi mod
// crtMod*auxMod with the factor split at 2^64 and 2^128 mod M folded in
// (uint128 arithmetic, wrapping where the reference's does)
u128 mult_qinvqi_aux(u64 input, u64 qinv, u64 qi) {
    const u128 M = (u128)Q * AUX;
    const u128 two128 = ((~(u128)0) % M + 1) % M;
    const u128 mask = ((u128)1 << 64) - 1;
    const u128 f = (u128)qi * qinv;
    const u128 temp = (u128)input * (f >> 64);
    u128 r = (temp >> 64) * two128 % M;
    r += (temp << 64) % M;
    r += (u128)input * (f & mask);
    return r % M;
}

// utils.cpp:535-553 crt_inv_bsgs_aux
inline u128 crt_inv_bsgs_aux(u64 x, u64 y) {
    const u128 M = (u128)Q * AUX;
    u128 r = mult_qinvqi_aux(x, AUX_INV_Q, AUX);
    r += mult_qinvqi_aux(y, Q_INV_AUX, Q);
    return r % M;
}

// utils.cpp:1191-1256 decompose_bsgs_aux (unsigned digits with offset, `d > M`)
void decompose_bsgs_aux(std::vector<Poly>& out_q, std::vector<Poly>& out_aux, const u64* in_q,
                        const u64* in_aux, int ell, u64 base, u64 bg) {
    const u128 M = (u128)Q * AUX;
    const int gbits = ilog2(bg);
    const int basebits = base != 0 ? ilog2(base) : 0;
    const int nsub = 128 - gbits;
    u128 offset = 0;
    for (int i = 0; i < ell; ++i) offset += (u128)1 << (basebits + i * gbits);
    offset <<= (gbits - 1);
    offset += base / 2;
    for (int i = 0; i < N; ++i) {
        u128 d = crt_inv_bsgs_aux(in_q[i], in_aux[i]) + offset;
        if (d > M) d -= M;
        d >>= basebits;
        for (int j = 0; j < ell; ++j) {
            i64 r = (i64)(u64)(d << nsub >> nsub);
            r -= (i64)(bg / 2);
            d >>= gbits;
            if (r >= 0) { out_q[j][i] = (u64)r; out_aux[j][i] = (u64)r; }
            else { out_q[j][i] = (u64)(r + (i64)Q); out_aux[j][i] = (u64)(r + (i64)AUX); }
        }
    }
}

// twosteps.cpp:1588-1715 evalAutoRNSKey, auxMod branch.  key: [2*ell][2][N],
// rows 0..ell-1 mod crtMod and ell..2ell-1 mod auxMod, NTT form.  The inputs
// come by value: the reference transforms them in place, on the copies its
// caller takes from keyMap.  Results are in coefficient form.
void eval_auto_rns_key(Rlwe& r1, Rlwe& r2, Rlwe in1, Rlwe in2, int32_t index, const u64* key,
                       int ell, u64 pp, u64 bg) {
    if (in1.isntt) { ntt_q().inverse(in1.a.data()); ntt_q().inverse(in1.b.data()); in1.isntt = false; }
    if (in2.isntt) { ntt_aux().inverse(in2.a.data()); ntt_aux().inverse(in2.b.data()); in2.isntt = false; }
    Poly ta1(N), tb1(N), ta2(N), tb2(N);
    automorphic_raw(ta1.data(), in1.a.data(), index, Q);
    automorphic_raw(tb1.data(), in1.b.data(), index, Q);
    automorphic_raw(ta2.data(), in2.a.data(), index, AUX);
    automorphic_raw(tb2.data(), in2.b.data(), index, AUX);
    std::vector<Poly> d1(ell, Poly(N)), d2(ell, Poly(N));
    decompose_bsgs_aux(d1, d2, ta1.data(), ta2.data(), ell, pp, bg);
    // crtMod part: (0, b) - g^-1(a) * key
    std::fill(r1.a.begin(), r1.a.end(), 0);
    r1.b = tb1;
    ntt_q().forward(r1.b.data());
    for (int i = 0; i < ell; ++i) {
        ntt_q().forward(d1[i].data());
        const u64* k = key + (size_t)i * 2 * N;
        ew_submul(r1.a.data(), d1[i].data(), k, mq());
        ew_submul(r1.b.data(), d1[i].data(), k + N, mq());
    }
    ntt_q().inverse(r1.a.data());
    ntt_q().inverse(r1.b.data());
    // auxMod part, digits transformed from their auxMod residues
    std::fill(r2.a.begin(), r2.a.end(), 0);
    r2.b = tb2;
    ntt_aux().forward(r2.b.data());
    for (int i = 0; i < ell; ++i) {
        ntt_aux().forward(d2[i].data());
        const u64* k = key + (size_t)(i + ell) * 2 * N;
        ew_submul(r2.a.data(), d2[i].data(), k, maux());
        ew_submul(r2.b.data(), d2[i].data(), k + N, maux());
    }
    ntt_aux().inverse(r2.a.data());
    ntt_aux().inverse(r2.b.data());
}

// twosteps.cpp:1814-1863, the giant-step half of genAutoKeyFromOffline.
//   auto2    [6][2][N]   keyMap.at(5^N1) of AutoKeyBSGSRNS(N, crtMod, auxMod, 3, 2^11*auxMod, 2^15), NTT form
//   offline2 [10][2][N]  keyMap.at(5^N1) of AutoKeyBSGSRNS(N, crtMod, auxMod, 5, 2^14, 2^14), NTT form
//   chain    [N2-2][6][2][N]  keys the reference inserts into result2 at 5^(k*N1), k = 2..N2-1
//                             (coefficient form), or NULL
//   giant    [N2-1][3][2][N]  result.keyMap at 5^(k*N1), k = 1..N2-1: auxMod^-1 * (x_crt - x_aux)
//                             mod crtMod, NTT form.  For k = 1 the inputs are the words of auto2
//                             as stored (NTT form), which the reference switches as they are.
void giant_from_offline(int n1, const u64* offline2, const u64* auto2, u64* chain, u64* giant) {
    const int n2 = N / 2 / n1, ell = 3;
    const int32_t step = pow_mod2(5, n1, 2 * N);
    std::vector<std::vector<Rlwe>> keys(1);
    for (int j = 0; j < 2 * ell; ++j) {
        Rlwe c(j < ell ? Q : AUX);
        const u64* src = auto2 + (size_t)j * 2 * N;
        c.a.assign(src, src + N);
        c.b.assign(src + N, src + 2 * N);
        c.isntt = true;
        keys[0].push_back(c);
    }
    for (int i = 1; i < n2 - 1; ++i) {
        std::vector<Rlwe> next;
        for (int j = 0; j < 2 * ell; ++j) next.emplace_back(j < ell ? Q : AUX);
        for (int j = 0; j < ell; ++j)
            eval_auto_rns_key(next[j], next[j + ell], keys.back()[j], keys.back()[j + ell], step, offline2, 5,
                              1ULL << 14, 1ULL << 14);
        keys.push_back(next);
        if (chain)
            for (int j = 0; j < 2 * ell; ++j) {
                u64* o = chain + (((size_t)(i - 1) * 2 * ell + j) * 2) * N;
                std::memcpy(o, next[j].a.data(), 8 * N);
                std::memcpy(o + N, next[j].b.data(), 8 * N);
            }
    }
    const Mod& m = mq();
    for (int i = 0; i < n2 - 1; ++i)
        for (int j = 0; j < ell; ++j)
            for (int part = 0; part < 2; ++part) {
                const Poly& x1 = part ? keys[i][j].b : keys[i][j].a;
                const Poly& x3 = part ? keys[i][j + ell].b : keys[i][j + ell].a;
                u64* o = giant + (((size_t)i * ell + j) * 2 + part) * N;
                for (int k = 0; k < N; ++k) o[k] = m.mul(m.sub(x1[k], x3[k]), AUX_INV_Q);
                ntt_q().forward(o);
            }
}

// encrypt.cpp:300-387 encrypt_rns_bsgs_autokey with modulus2 = auxMod
void encrypt_rns_aux_autokey(u64* b1, u64* a1, u64* b2, u64* a2, const Secret& s, const Poly& msg1,
                             const Poly& msg2) {
    Poly s1 = s.coeff, s2 = s.coeff;
    guass_to_modulus(s2.data(), Q, AUX);
    ntt_q().forward(s1.data());
    ntt_aux().forward(s2.data());
    const Mod &m1 = mq(), &m2 = maux();
    sample_random(a1, Q, N);
    sample_random(a2, AUX, N);
    ew_mul(b1, a1, s1.data(), m1);
    ew_mul(b2, a2, s2.data(), m2);
    Poly err(N), temp(N);
    sample_guass(err.data(), Q);
    ew_add(temp.data(), err.data(), msg1.data(), m1);
    ntt_q().forward(temp.data());
    ew_add(b1, b1, temp.data(), m1);
    guass_to_modulus(err.data(), Q, AUX);
    ew_add(temp.data(), err.data(), msg2.data(), m2);
    ntt_aux().forward(temp.data());
    ew_add(b2, b2, temp.data(), m2);
}

// lwe.cpp:1025-1070 generateSingleKeyBS for AutoKeyBSGSRNS(N, crtMod, auxMod, ell, pp, bg):
// out [2*ell][2][N], NTT form
void aux_key(u64* out, int32_t index, const Secret& s, int ell, u64 pp, u64 bg) {
    Poly autosec(N), autosec2(N), t1(N), t2(N);
    std::vector<u128> mf = power_of_bg128(pp, bg, ell);
    automorphic(autosec.data(), s.coeff.data(), index, Q);
    autosec2 = autosec;
    guass_to_modulus(autosec2.data(), Q, AUX);
    for (int i = 0; i < ell; ++i) {
        ew_mulc(t1.data(), autosec.data(), (u64)(mf[i] % Q), mq());
        ew_mulc(t2.data(), autosec2.data(), (u64)(mf[i] % AUX), maux());
        u64* kq = out + (size_t)i * 2 * N;
        u64* ka = out + (size_t)(i + ell) * 2 * N;
        encrypt_rns_aux_autokey(kq + N, kq, ka + N, ka, s, t1, t2);
    }
}

double now_us() {
    return std::chrono::duration<double, std::micro>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

// utils.cpp:118-203 QInv(lwenum, crtMod)
u64 qinv_table(int r) {
    switch (r) {
        case 1: return 1ULL;
        case 2: return 33487344869801985ULL;
        case 4: return 50231017304702977ULL;
        case 8: return 58602853522153473ULL;
        case 16: return 62788771630878721ULL;
        case 32: return 64881730685241345ULL;
        case 64: return 65928210212422657ULL;
        case 256: return 66713069857808641ULL;
        case 512: return 66843879798706305ULL;
        case 4096: return 66958338496991761ULL;
        default: return 0;
    }
}

bool valid_shape(int r, int n1) {
    if (r < 1 || (r & (r - 1)) || r > N) return false;
    if (n1 < 2 || (n1 & (n1 - 1)) || n1 > N / 2) return false;
    return true;
}

}  // namespace

// =================================================================== C API

extern "C" {

uint64_t ko_const(int which) {
    switch (which) {
        case 0: return N;
        case 1: return Q1;
        case 2: return Q2;
        case 3: return Q;
        case 4: return BS;
        case 5: return PT;
        case 6: return DELTA;
        case 7: return ROOT_Q;
        case 8: return ntt_bs().root;
        case 9: return ntt_p().root;
        case 10: return AUX;
        case 11: return AUX_INV_Q;
        case 12: return Q_INV_AUX;
        case 13: return ntt_aux().root;
        default: return 0;
    }
}

uint64_t ko_qinv(int r) { return qinv_table(r); }

void ko_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

static const Ntt& ntt_sel(int modsel) {
    static Ntt t1(Q1, ROOT_Q % Q1), t2(Q2, ROOT_Q % Q2);
    switch (modsel) {
        case 0: return ntt_q();
        case 1: return ntt_bs();
        case 2: return ntt_p();
        case 3: return t1;
        case 5: return ntt_aux();
        default: return t2;
    }
}

void ko_ntt_forward(int modsel, uint64_t* a) { ntt_sel(modsel).forward(a); }
int ko_ntt_set_vector(int on) {
    g_ntt_vector = on != 0;
    return ntt_vector_usable() ? 1 : 0;
}
void ko_ntt_inverse(int modsel, uint64_t* a) { ntt_sel(modsel).inverse(a); }
uint64_t ko_min_root(uint64_t q) { return min_root(q); }
uint64_t ko_crt_compose(uint64_t x, uint64_t y) { return crt_compose(x, y); }

void ko_decompose(uint64_t* out, const uint64_t* in, int ell, uint64_t base, uint64_t bg, uint64_t q) {
    std::vector<Poly> d(ell, Poly(N));
    decompose(d, in, ell, base, bg, q);
    for (int j = 0; j < ell; ++j) std::memcpy(out + (size_t)j * N, d[j].data(), N * 8);
}

void ko_decompose_rlwe(uint64_t* out, const uint64_t* in_a, const uint64_t* in_b, int ell,
                       uint64_t base, uint64_t bg, uint64_t q) {
    std::vector<Poly> d(2 * ell, Poly(N));
    decompose_rlwe(d, in_a, in_b, ell, base, bg, q);
    for (int j = 0; j < 2 * ell; ++j) std::memcpy(out + (size_t)j * N, d[j].data(), N * 8);
}

void ko_decompose_bsgs(uint64_t* out_q, uint64_t* out_bs, const uint64_t* in_q, const uint64_t* in_bs,
                       int ell, uint64_t base, uint64_t bg) {
    std::vector<Poly> dq(ell, Poly(N)), db(ell, Poly(N));
    decompose_bsgs(dq, db, in_q, in_bs, ell, base, bg);
    for (int j = 0; j < ell; ++j) {
        std::memcpy(out_q + (size_t)j * N, dq[j].data(), N * 8);
        std::memcpy(out_bs + (size_t)j * N, db[j].data(), N * 8);
    }
}

void ko_crt_inv_bsgs(uint64_t* out, const uint64_t* in_q, const uint64_t* in_bs) {
    for (int i = 0; i < N; ++i) {
        u128 v = crt_inv_bsgs(in_q[i], in_bs[i]);
        out[2 * i] = (u64)v;
        out[2 * i + 1] = (u64)(v >> 64);
    }
}

void ko_hexl_ntt_index(int32_t* out) {
    std::vector<int32_t> h, rot;
    hexl_rotate_indexes(h, rot);
    std::memcpy(out, h.data(), N * 4);
}

void ko_query_encode(int32_t* out) {
    std::vector<int32_t> e = query_encode();
    std::memcpy(out, e.data(), N * 4);
}

void ko_permutation_matrix(int32_t* out, int n1) {
    std::vector<int32_t> p = permutation_matrix(n1);
    std::memcpy(out, p.data(), p.size() * 4);
}

void ko_automorphic_rns(uint64_t* out, const uint64_t* in, int index, uint64_t q) {
    automorphic_rns(out, in, index, q);
}

void ko_db_product_scalar(uint64_t* out, const uint64_t* db, const uint64_t* cbuf, int n1, int n2) {
    db_product_scalar(out, db, cbuf, n1, n2);
}
int ko_db_product_vectorised(void) { return __builtin_cpu_supports("avx512f") ? 1 : 0; }

void ko_db_product(uint64_t* out, const uint64_t* db, const uint64_t* cbuf, int n1, int n2) {
    db_product(out, db, cbuf, n1, n2);
}

int ko_gen_db(int r, int n1, unsigned seed, int plant, int tcol, int tpack, uint64_t* db) {
    if (!valid_shape(r, n1)) return -1;
    srand(seed);
    std::vector<u64> data((size_t)N * (N / 2)), data_ntt((size_t)(N / 2) * N);
    const size_t words = (size_t)N * N / 2;
    for (int k = 0; k < r; ++k) make_db_block(db + words * k, n1, k, r, plant, tcol, tpack, data, data_ntt);
    return 0;
}

// ko_gen_db with the records of each block kept as well: records [r][N][N/2]
// after sampling and planting (the input of database_tobsgsntt), db the
// database_tocrt words
int ko_gen_records(int r, int n1, unsigned seed, int plant, int tcol, int tpack, uint64_t* records, uint64_t* db) {
    if (!valid_shape(r, n1)) return -1;
    srand(seed);
    std::vector<u64> data((size_t)N * (N / 2)), data_ntt((size_t)(N / 2) * N);
    const size_t words = (size_t)N * N / 2;
    for (int k = 0; k < r; ++k) {
        make_db_block(db + words * k, n1, k, r, plant, tcol, tpack, data, data_ntt);
        std::memcpy(records + words * k, data.data(), words * 8);
    }
    return 0;
}

// inverse_encode (twosteps.cpp:182-198) of one record column of N values, in place
void ko_inverse_encode(uint64_t* col) {
    Poly m(col, col + N);
    inverse_encode(m);
    std::memcpy(col, m.data(), (size_t)N * 8);
}

// database_tobsgsntt + database_tocrt (twosteps.cpp:42-179, crt.cpp:859-885)
// of r blocks of records [N][N/2], block k at records + k * stride (stride 0
// repeats block 0): words [r][N*N/2], as the preprocessing of test-pir.cpp:46-63
int ko_records_to_words(int r, int n1, const uint64_t* records, int64_t stride, uint64_t* words) {
    if (!valid_shape(r, n1) || stride < 0) return -1;
    std::vector<u64> data((size_t)N * (N / 2)), data_ntt((size_t)(N / 2) * N);
    const size_t w = (size_t)N * N / 2;
    for (int k = 0; k < r; ++k) {
        std::memcpy(data.data(), records + (size_t)stride * k, w * 8);
        database_tobsgsntt(data_ntt, data, n1);
        database_tocrt(words + w * k, data_ntt, n1);
    }
    return 0;
}

// database_tobsgsntt alone on one block: data [N][N/2] -> result [N/2][N]
int ko_database_tobsgsntt(int n1, const uint64_t* data, uint64_t* result) {
    if (n1 < 1 || n1 > N / 2 || (n1 & (n1 - 1))) return -1;
    std::vector<u64> d(data, data + (size_t)N * N / 2), res((size_t)(N / 2) * N);
    database_tobsgsntt(res, d, n1);
    std::memcpy(result, res.data(), res.size() * 8);
    return 0;
}

static void gen_client_tail(int r, int n1, int tcol, int tpack, const Secret& s, uint64_t* query,
                            uint64_t* rgsw, uint64_t* baby, uint64_t* giant, uint64_t* packing) {
    // test-pir.cpp:68-89: query, RGSW, BSGS keys, packing keys
    query_bsgs_rns(query, s, (u64)tcol);
    rgsw_keygen(rgsw, s, (u64)tpack);
    bsgs_keygen(baby, giant, s, n1);
    packing_keygen(packing, s, r);
}

int ko_gen_client(int r, int n1, unsigned seed, int tcol, int tpack, uint64_t* secret, uint64_t* query,
                  uint64_t* rgsw, uint64_t* baby, uint64_t* giant, uint64_t* packing) {
    if (!valid_shape(r, n1) || tcol < 0 || tcol >= N / 2 || tpack < 0 || tpack >= r) return -1;
    srand(seed);
    Secret s = make_secret();
    std::memcpy(secret, s.coeff.data(), N * 8);
    gen_client_tail(r, n1, tcol, tpack, s, query, rgsw, baby, giant, packing);
    return 0;
}

int ko_gen_instance(int r, int n1, unsigned seed, int plant, uint64_t* secret, int32_t* target,
                    uint64_t* db, uint64_t* query, uint64_t* rgsw, uint64_t* baby, uint64_t* giant,
                    uint64_t* packing, int32_t* perms) {
    if (!valid_shape(r, n1)) return -1;
    srand(seed);
    // test-pir.cpp:20-24
    Secret s = make_secret();
    std::memcpy(secret, s.coeff.data(), N * 8);
    int32_t tcol = rand() % (N / 2);
    int32_t tpack = rand() % r;
    target[0] = tcol;
    target[1] = tpack;
    // test-pir.cpp:46-63
    std::vector<u64> data((size_t)N * (N / 2)), data_ntt((size_t)(N / 2) * N);
    const size_t words = (size_t)N * N / 2;
    for (int k = 0; k < r; ++k) make_db_block(db + words * k, n1, k, r, plant, tcol, tpack, data, data_ntt);
    gen_client_tail(r, n1, tcol, tpack, s, query, rgsw, baby, giant, packing);
    // test-pir.cpp:97-98
    std::vector<int32_t> p = permutation_matrix(n1);
    std::memcpy(perms, p.data(), p.size() * 4);
    return 0;
}

int ko_answer(int r, int n1, const uint64_t* db, const uint64_t* query, const uint64_t* rgsw,
              const uint64_t* baby, const uint64_t* giant, const uint64_t* packing,
              const int32_t* perms, uint64_t* response, uint64_t* cipherbuf_out,
              uint64_t* firstdim, uint64_t* ext_out) {
    if (!valid_shape(r, n1) || n1 < 4) return -1;
    const int n2 = N / 2 / n1;
    const size_t words = (size_t)N * N / 2;
    for (double& v : g_stage_us) v = 0;
    double t0 = now_us();

    // matrix_vector_mul_bsgs_rns_crt_large (twosteps.cpp:1527-1579)
    Rlwe input[2] = {Rlwe(Q), Rlwe(BS)};
    std::memcpy(input[0].a.data(), query + 0 * N, N * 8);
    std::memcpy(input[0].b.data(), query + 1 * N, N * 8);
    std::memcpy(input[1].a.data(), query + 2 * N, N * 8);
    std::memcpy(input[1].b.data(), query + 3 * N, N * 8);
    std::vector<Rlwe> rot(n1 - 1, Rlwe(Q));
    {
        // evalAutoSRNSSmallKeys (twosteps.cpp:1206-1214)
        Rlwe inter[2] = {Rlwe(Q), Rlwe(BS)};
        eval_auto_srns_first(rot, inter, input, n1 / 2, baby, perms);
        eval_auto_srns_second(rot, inter, n1 / 2, baby, perms);
    }
    std::vector<u64> cbuf((size_t)n1 * N * 2);
    reorient_cipher(cbuf.data(), input[0], rot, n1);
    if (cipherbuf_out) std::memcpy(cipherbuf_out, cbuf.data(), cbuf.size() * 8);
    double t1 = now_us();
    g_stage_us[0] = t1 - t0;

    std::vector<Rlwe> result(r, Rlwe(Q));
    std::vector<u64> out((size_t)n2 * 2 * N);
    for (int i = 0; i < r; ++i) {
        double a = now_us();
        db_product(out.data(), db + words * i, cbuf.data(), n1, n2);
        double b = now_us();
        std::memcpy(result[i].a.data(), out.data(), N * 8);
        std::memcpy(result[i].b.data(), out.data() + N, N * 8);
        result[i].isntt = true;
        eval_auto_rns_crt(result[i], out.data(), n2, giant);
        double c = now_us();
        g_stage_us[1] += b - a;
        g_stage_us[2] += c - b;
    }
    if (firstdim)
        for (int i = 0; i < r; ++i) {
            std::memcpy(firstdim + (size_t)i * 2 * N, result[i].a.data(), N * 8);
            std::memcpy(firstdim + (size_t)i * 2 * N + N, result[i].b.data(), N * 8);
        }

    // test-pir.cpp:114-115
    double t2 = now_us();
    std::vector<Rlwe> ext(r, Rlwe(Q));
    for (int j = 0; j < r; ++j) external_product(ext[j], result[j], rgsw);
    if (ext_out)
        for (int i = 0; i < r; ++i) {
            std::memcpy(ext_out + (size_t)i * 2 * N, ext[i].a.data(), N * 8);
            std::memcpy(ext_out + (size_t)i * 2 * N + N, ext[i].b.data(), N * 8);
        }
    double t3 = now_us();
    g_stage_us[3] = t3 - t2;

    Rlwe res(Q);
    packing_rlwes(res, ext, packing);
    std::memcpy(response, res.a.data(), N * 8);
    std::memcpy(response + N, res.b.data(), N * 8);
    g_stage_us[4] = now_us() - t3;
    return 0;
}

// `count` independent online answers of the same inputs on `threads` OpenMP
// threads: the all-cores CPU baseline of a batch.  Returns 0 when every
// response equals the first (written to `response`), -1 on a shape error,
// -2 on a mismatch.
int ko_answer_parallel(int count, int threads, int r, int n1, const uint64_t* db, const uint64_t* query,
                       const uint64_t* rgsw, const uint64_t* baby, const uint64_t* giant,
                       const uint64_t* packing, const int32_t* perms, uint64_t* response) {
    if (count < 1 || threads < 1 || !valid_shape(r, n1) || n1 < 4) return -1;
    std::vector<std::vector<u64>> outs(count, std::vector<u64>(2 * N));
    int bad = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic) reduction(+ : bad)
    for (int i = 0; i < count; ++i)
        bad += ko_answer(r, n1, db, query, rgsw, baby, giant, packing, perms, outs[i].data(), nullptr, nullptr,
                         nullptr) != 0;
    if (bad) return -1;
    for (int i = 1; i < count; ++i)
        if (outs[i] != outs[0]) return -2;
    std::memcpy(response, outs[0].data(), 2 * N * 8);
    return 0;
}

void ko_last_stage_us(double* out5) {
    for (int i = 0; i < 5; ++i) out5[i] = g_stage_us[i];
}

int ko_decrypt(int r, const uint64_t* secret, const uint64_t* response, uint64_t* msg) {
    // twosteps.cpp:388-418 decrypt_bsgs_total; encrypt.cpp:528-573 decrypt
    const Mod& m = mq();
    Poly a(response, response + N), b(response + N, response + 2 * N), s(secret, secret + N);
    ntt_q().forward(a.data());
    ntt_q().forward(b.data());
    ntt_q().forward(s.data());
    Poly mm(N);
    ew_mul(mm.data(), a.data(), s.data(), m);
    ew_sub(mm.data(), b.data(), mm.data(), m);
    ntt_q().inverse(mm.data());
    ew_mulc(mm.data(), mm.data(), qinv_table(r), m);
    for (int i = 0; i < N; ++i) {
        u64 me = mm[i];
        if (me > Q / 2) me = (u64)(roundl(((long double)me - Q) / DELTA) + PT);
        else me = (u64)roundl((long double)me / DELTA);
        msg[i] = me;
    }
    return 0;
}

void ko_crt_inv_bsgs_aux(uint64_t* out, const uint64_t* in_q, const uint64_t* in_aux) {
    for (int i = 0; i < N; ++i) {
        u128 v = crt_inv_bsgs_aux(in_q[i], in_aux[i]);
        out[2 * i] = (u64)v;
        out[2 * i + 1] = (u64)(v >> 64);
    }
}

void ko_decompose_bsgs_aux(uint64_t* out_q, uint64_t* out_aux, const uint64_t* in_q, const uint64_t* in_aux,
                           int ell, uint64_t base, uint64_t bg) {
    std::vector<Poly> oq(ell, Poly(N)), oa(ell, Poly(N));
    decompose_bsgs_aux(oq, oa, in_q, in_aux, ell, base, bg);
    for (int j = 0; j < ell; ++j) {
        std::memcpy(out_q + (size_t)j * N, oq[j].data(), 8 * N);
        std::memcpy(out_aux + (size_t)j * N, oa[j].data(), 8 * N);
    }
}

int ko_gen_offline_giant(int n1, unsigned seed, uint64_t* secret, uint64_t* offline2, uint64_t* auto2) {
    if (!valid_shape(1, n1) || N / 2 / n1 < 2) return -1;
    srand(seed);
    Secret s = make_secret();
    const int32_t index = pow_mod2(5, n1, 2 * N);
    aux_key(offline2, index, s, 5, 1ULL << 14, 1ULL << 14);
    aux_key(auto2, index, s, 3, (1ULL << 11) * AUX, 1ULL << 15);
    std::memcpy(secret, s.coeff.data(), 8 * N);
    return 0;
}

int ko_giant_from_offline(int n1, const uint64_t* offline2, const uint64_t* auto2, uint64_t* chain,
                          uint64_t* giant) {
    if (!valid_shape(1, n1) || N / 2 / n1 < 2) return -1;
    giant_from_offline(n1, offline2, auto2, chain, giant);
    return 0;
}

int ko_rlwe_phase(int modsel, const uint64_t* secret, const uint64_t* a, const uint64_t* b, int ntt_form,
                  uint64_t* out) {
    // b - a*s in coefficient form, mod crtMod (modsel 0) or auxMod (modsel 5);
    // the secret is given mod crtMod (coefficients 0, 1, crtMod-1)
    if (modsel != 0 && modsel != 5) return -1;
    const u64 q = modsel ? AUX : Q;
    const Ntt& t = ntt_sel(modsel);
    const Mod m(q);
    Poly s(secret, secret + N), x(a, a + N), y(b, b + N);
    if (modsel) guass_to_modulus(s.data(), Q, AUX);
    t.forward(s.data());
    if (!ntt_form) { t.forward(x.data()); t.forward(y.data()); }
    for (int i = 0; i < N; ++i) y[i] = m.sub(y[i], m.mul(x[i], s[i]));
    t.inverse(y.data());
    std::memcpy(out, y.data(), 8 * N);
    return 0;
}

}  // extern "C"
