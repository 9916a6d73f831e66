// kspir_device.h — device-side arithmetic for the KsPIR online answer on
// CDNA4 (gfx950): exact modular arithmetic on the two 28-bit CRT primes of
// the ciphertext modulus Q = q1*q2 and the 24-bit baby-step prime bsMod, all
// in 32-bit lanes (mul_lo / mul_hi), plus a whole-polynomial negacyclic NTT
// staged in LDS.
//
// Residue number system: a value x mod Q is held as (x mod q1, x mod q2); a
// baby-step value additionally as x mod bsMod.  Every product is either
//   * Shoup:      x*w mod q with a precomputed w' = floor(w*2^32/q)
//                 (NTT twiddles and scalar constants), or
//   * Montgomery: lazily accumulated sum of up to 4 products of a residue with
//                 a key word stored as key*2^32 mod q, reduced once.
// Both return canonical residues in [0, q).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kspir {

constexpr int kN = 4096;       // ring dimension (params.h:16)
constexpr int kLogN = 12;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int kNttThreads = 256;   // threads per workgroup of every NTT kernel

// the three primes of the residue system (crt.h:16-17, :32)
constexpr uint32_t kQ1 = 268369921u;
constexpr uint32_t kQ2 = 249561089u;
constexpr uint32_t kBs = 16760833u;
constexpr uint32_t kP = 65537u;          // plaintext modulus (params.h)
constexpr uint32_t kAux = 268361729u;    // auxiliary prime of offline key derivation (crt.h:37)
constexpr uint64_t kQ = (uint64_t)kQ1 * kQ2;

enum Prime { P_Q1 = 0, P_Q2 = 1, P_BS = 2 };

// per-prime constants and twiddle tables resident in HBM
struct PrimeTab {
    uint32_t q;
    uint32_t mont_qinv;    // -q^-1 mod 2^32
    uint32_t ninv, ninv_sh;
    const uint32_t* fw;    // fw[bitrev(i)] = psi^i
    const uint32_t* fw_sh;
    const uint32_t* iv;    // iv[bitrev(i)] = psi^-i
    const uint32_t* iv_sh;
    const uint32_t* psi_pow;   // psi^e, e in [0, 2N)
    const uint2* fw2;          // (2^32 - fw[i], fw_sh[i]) pairs (forward butterflies use -w)
    const uint2* iv2;          // (iv[i], iv_sh[i]) pairs
};

struct Tabs {
    PrimeTab p[5];              // q1, q2, bsMod, plaintext p, auxMod
    const uint16_t* slot_exp;   // NTT slot s holds psi^(slot_exp[s]) (odd exponent)
    const uint16_t* exp_slot;   // exp_slot[e >> 1] = slot holding psi^e
    // CRT / modulus-switch constants (Shoup pairs)
    uint32_t q1inv_mod_q2, q1inv_mod_q2_sh;       // q1^-1 mod q2
    uint32_t bsinv_q[2], bsinv_q_sh[2];          // bsMod^-1 mod q1, q2
    uint32_t qinv_bs, qinv_bs_sh;                // Q^-1 mod bsMod
    uint32_t q_mod_bs;                           // Q mod bsMod
};

__device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) { return __umulhi(a, b); }

// The conditional corrections below are written as unsigned minima: for
// r < 2^32, min(r, r - q) (wrapping subtraction) equals r >= q ? r - q : r,
// and it compiles to a subtract and a v_min instead of a compare, a
// lane-mask select and the wait state the select needs after the compare.
__device__ __forceinline__ uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }
// a + b mod q for a, b < q
__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b, uint32_t q) {
    uint32_t s = a + b;
    return umin32(s, s - q);
}
// a - b mod q for a, b < q < 2^31: a - b if it did not wrap, else a - b + q
__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t q) {
    uint32_t d = a - b;
    return umin32(d, d + q);
}
// x * w mod q, x < 2^32, w < q, wp = floor(w * 2^32 / q)
__device__ __forceinline__ uint32_t mul_shoup(uint32_t x, uint32_t w, uint32_t wp, uint32_t q) {
    uint32_t qh = mulhi32(x, wp);
    uint32_t r = x * w - qh * q;
    return umin32(r, r - q);
}
// T * 2^-32 mod q for T < q * 2^32
__device__ __forceinline__ uint32_t mont_reduce(uint64_t T, uint32_t q, uint32_t qinv) {
    uint32_t m = (uint32_t)T * qinv;
    uint64_t t = T + (uint64_t)m * q;
    uint32_t r = (uint32_t)(t >> 32);
    return umin32(r, r - q);
}
// -a mod q for a < q: q - a, and 0 for a = 0 (q - 0 = q reduces to 0)
__device__ __forceinline__ uint32_t neg_mod(uint32_t a, uint32_t q) {
    uint32_t t = q - a;
    return umin32(t, t - q);
}

// value in [0, Q) from its residues (Garner): x1 + q1 * ((x2 - x1) * q1^-1 mod q2)
__device__ __forceinline__ uint64_t crt_q(uint32_t x1, uint32_t x2, const Tabs& T) {
    uint32_t x1m = x1 >= kQ2 ? x1 - kQ2 : x1;
    uint32_t t = mul_shoup(sub_mod(x2, x1m, kQ2), T.q1inv_mod_q2, T.q1inv_mod_q2_sh, kQ2);
    return (uint64_t)x1 + (uint64_t)kQ1 * t;
}

// residue mod q of a signed gadget digit r with |r| < q
__host__ __device__ __forceinline__ uint32_t digit_residue(int32_t r, uint32_t q) {
    // |r| < q < 2^31: u + q wraps to r + q exactly when r < 0, and is then the
    // smaller of the two; for r >= 0 it stays above u
    const uint32_t u = (uint32_t)r, w = u + q;
    return w < u ? w : u;
}

// ------------------------------------------------------ gadget digits

// utils.cpp:1120-1189 decompose_bsgs digit j (ell 3, PP = Bg = 2^20).  The
// digit helpers below are host-callable as well, so that their boundary cases
// can be checked against the oracle on the CPU.
__host__ __device__ __forceinline__ int32_t bsgs_digit(unsigned __int128 v, int j) {
    const unsigned __int128 M = (unsigned __int128)kQ * kBs;
    const unsigned __int128 offset = (((unsigned __int128)1 << 20) + ((unsigned __int128)1 << 40) +
                                      ((unsigned __int128)1 << 60)) << 19 | ((unsigned __int128)1 << 19);
    unsigned __int128 d = v + offset;
    if (d > M) d -= M;
    d >>= 20;
    uint32_t r = (uint32_t)(d >> (20 * j)) & ((1u << 20) - 1);
    return (int32_t)r - (1 << 19);
}

// utils.cpp:754-811 decompose (tfhe variant): digit j of x in [0, Q]
template <int GBITS, int BASEBITS, int ELL>
__host__ __device__ __forceinline__ int32_t tfhe_digit(uint64_t x, int j) {
    uint64_t offset = 0;
#pragma unroll
    for (int i = 0; i < ELL; ++i) offset += (uint64_t)1 << (BASEBITS + i * GBITS);
    offset <<= (GBITS - 1);
    offset += BASEBITS ? ((uint64_t)1 << (BASEBITS - 1)) : 0;
    uint64_t d = x + offset;
    if (d > kQ) d -= kQ;
    d >>= BASEBITS;
    uint32_t r = (uint32_t)(d >> (GBITS * j)) & ((1u << GBITS) - 1);
    return (int32_t)r - (1 << (GBITS - 1));
}

// utils.cpp:974-1044 decompose_rlwe (signed digits): digit j of x in [0, Q)
template <int GBITS, int BASEBITS>
__host__ __device__ __forceinline__ int32_t signed_digit(uint64_t x, int j) {
    int64_t d = x > (kQ >> 1) ? (int64_t)(x - kQ) : (int64_t)x;
    d >>= BASEBITS;
    int32_t r = 0;
    for (int k = 0; k <= j; ++k) {
        r = (int32_t)((int64_t)((uint64_t)d << (64 - GBITS)) >> (64 - GBITS));
        d -= r;
        d >>= GBITS;
    }
    return r;
}

// ---------------------------------------------------------------- NTT in LDS
// s: 4096 residues in LDS.  Forward: Cooley-Tukey over fw[m + i] (natural
// order in, slot order out: slot i = a(psi^(2*bitrev(i)+1))).  Inverse:
// Gentleman-Sande over iv[m + i], then scaled by N^-1.  All values stay in
// [0, q).  Caller synchronises before and after.
__device__ __forceinline__ void ntt_fwd_lds(uint32_t* s, const PrimeTab& P) {
    const uint32_t q = P.q;
    const int tid = threadIdx.x;
#pragma unroll 1
    for (int lg = 0; lg < kLogN; ++lg) {
        const int m = 1 << lg;
        const int lt = kLogN - 1 - lg;    // log2 t
        const int t = 1 << lt;
#pragma unroll
        for (int k = 0; k < kN / 2 / kNttThreads; ++k) {
            const int b = tid + k * kNttThreads;
            const int i = b >> lt;
            const int j = b & (t - 1);
            const int x = (i << (lt + 1)) + j;
            const uint32_t w = P.fw[m + i], wp = P.fw_sh[m + i];
            uint32_t u = s[x];
            uint32_t v = mul_shoup(s[x + t], w, wp, q);
            s[x] = add_mod(u, v, q);
            s[x + t] = sub_mod(u, v, q);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void ntt_inv_lds(uint32_t* s, const PrimeTab& P) {
    const uint32_t q = P.q;
    const int tid = threadIdx.x;
#pragma unroll 1
    for (int lg = kLogN - 1; lg >= 0; --lg) {
        const int m = 1 << lg;
        const int lt = kLogN - 1 - lg;
        const int t = 1 << lt;
#pragma unroll
        for (int k = 0; k < kN / 2 / kNttThreads; ++k) {
            const int b = tid + k * kNttThreads;
            const int i = b >> lt;
            const int j = b & (t - 1);
            const int x = (i << (lt + 1)) + j;
            const uint32_t w = P.iv[m + i], wp = P.iv_sh[m + i];
            uint32_t u = s[x], v = s[x + t];
            s[x] = add_mod(u, v, q);
            s[x + t] = mul_shoup(u + q - v, w, wp, q);
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kN / kNttThreads; ++k) {
        const int x = threadIdx.x + k * kNttThreads;
        s[x] = mul_shoup(s[x], P.ninv, P.ninv_sh, q);
    }
    __syncthreads();
}

// ------------------------------------------------- register-blocked NTT
// The same transforms as ntt_fwd_lds / ntt_inv_lds, organised as three
// passes of four radix-2 stages.  In pass P a thread holds 16 elements whose
// indices differ only in bits 4P..4P+3 reversed into its register index k,
// so the four stages of a pass run in registers and the passes exchange
// through LDS: two barriers per transform instead of twelve.
//   pass 0: e = tid + 256k
//   pass 1: e = 256*g + (tid & 15) + 16k,  g = pass_base(1, tid)
//   pass 2: e = 16*c + k,                  c = pass_base(2, tid)
// LDS words carry one pad word per 64 (kPadN words).  With that padding a
// wave's 64 lanes hit 64 distinct banks in every pass (and each half-wave
// 32 distinct banks of 32) because pass 1 puts the groups w, w+4, w+8,
// w+12 in wave w and pass 2 orders a wave's columns so that each half-wave
// holds the columns c with c & 2 fixed; consecutive-index loops
// (x = tid + 256i) are conflict-free as well.
constexpr int kPadN = kN + kN / 64;
__host__ __device__ __forceinline__ int lds_pad(int e) { return e + (e >> 6); }

// pass 1: group g (its 256-element block); pass 2: column c; pass 0: 0
__host__ __device__ __forceinline__ int pass_base(int pass, int tid) {
    return pass == 0 ? 0
         : pass == 1 ? ((((tid >> 4) & 3) << 2) | (tid >> 6))
                     : ((tid & ~63) | (tid & 1) | (((tid >> 5) & 1) << 1) | (((tid >> 1) & 15) << 2));
}

__host__ __device__ __forceinline__ int pass_elem(int pass, int tid, int k) {
    return pass == 0 ? tid + 256 * k
         : pass == 1 ? (pass_base(1, tid) << 8) + (tid & 15) + 16 * k
                     : 16 * pass_base(2, tid) + k;
}

__host__ __device__ __forceinline__ uint32_t mulhi32_hd(uint32_t a, uint32_t b) {
    return (uint32_t)(((uint64_t)a * b) >> 32);
}
__host__ __device__ __forceinline__ uint32_t mul_shoup_hd(uint32_t x, uint32_t w, uint32_t wp, uint32_t q) {
    uint32_t qh = mulhi32_hd(x, wp);
    uint32_t r = x * w - qh * q;
    return r >= q ? r - q : r;
}

// Twiddles of pass P for thread tid (forward tables hold 2^32 - w): stage s (s = 0..3) of the pass uses the
// 2^s consecutive table entries starting at 2^(4P+s) + base * 2^s, with
// base = pass_base(P, tid), i.e. the entry for register k is
// table[2^(4P+s) + base * 2^s + (k >> (4 - s))].  tw2 interleaves (w, w')
// pairs so each stage's group is one aligned vector load; w[], ws[] hold
// the 15 pairs as stage s at offset 2^s - 1.
__host__ __device__ __forceinline__ void load_pass_twiddles(const uint2* tw2, int pass, int tid, uint32_t w[15],
                                                            uint32_t ws[15]) {
    const int base = pass_base(pass, tid);
    const int lg0 = 4 * pass;
    // s = 0: one pair
    {
        const uint2 t = tw2[(1 << lg0) + base];
        w[0] = t.x; ws[0] = t.y;
    }
    // s = 1: two pairs (16 bytes)
    {
        const uint4 t = *reinterpret_cast<const uint4*>(tw2 + (1 << (lg0 + 1)) + base * 2);
        w[1] = t.x; ws[1] = t.y; w[2] = t.z; ws[2] = t.w;
    }
    // s = 2: four pairs
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint4 t = *reinterpret_cast<const uint4*>(tw2 + (1 << (lg0 + 2)) + base * 4 + 2 * h);
        w[3 + 2 * h] = t.x; ws[3 + 2 * h] = t.y; w[4 + 2 * h] = t.z; ws[4 + 2 * h] = t.w;
    }
    // s = 3: eight pairs
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const uint4 t = *reinterpret_cast<const uint4*>(tw2 + (1 << (lg0 + 3)) + base * 8 + 2 * h);
        w[7 + 2 * h] = t.x; ws[7 + 2 * h] = t.y; w[8 + 2 * h] = t.z; ws[8 + 2 * h] = t.w;
    }
}

__host__ __device__ __forceinline__ uint32_t umin_hd(uint32_t a, uint32_t b) { return a < b ? a : b; }

// 64-bit a * b and a * b + c (the callers use the low words, which the
// compiler forms with 32-bit multiplies and adds)
__host__ __device__ __forceinline__ uint64_t mul_wide(uint32_t a, uint32_t b) { return (uint64_t)a * b; }
__host__ __device__ __forceinline__ uint64_t mad_wide(uint32_t a, uint32_t b, uint64_t c) { return (uint64_t)a * b + c; }

// x * w mod q up to one extra q: result in [0, 2q) for any x < 2^32 (Shoup,
// without the final correction), w < q, wp = floor(w * 2^32 / q):
// x*w - floor(x*wp / 2^32)*q, formed as mulhi * (2^32 - q) + x*w
__host__ __device__ __forceinline__ uint32_t mul_shoup_lazy(uint32_t x, uint32_t w, uint32_t wp, uint32_t q) {
    return (uint32_t)mad_wide(mulhi32_hd(x, wp), 0u - q, mul_wide(x, w));
}

// Butterflies with lazy reduction (Harvey, "Faster arithmetic for
// number-theoretic transforms", 2014): every q here is below 2^28, so words
// up to 4q fit in 32 bits and the butterflies skip most corrections.
// Forward pass (stages 4P .. 4P+3, Cooley-Tukey) on registers v[16]:
// inputs in [0, 4q), outputs in [0, 4q).
__host__ __device__ __forceinline__ void ntt_fwd_pass(uint32_t v[16], const uint32_t w[15], const uint32_t ws[15],
                                                      uint32_t q) {
    const uint32_t q2 = 2 * q;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const int kb = 8 >> st;   // register-index bit flipped by this stage
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (k & kb) continue;
            const int ti = (1 << st) - 1 + (k >> (4 - st));
            const uint32_t u = umin_hd(v[k], v[k] - q2);                    // [0, 2q)
            // nt = -t mod 2^32 for t = w * v[k|kb] mod q in [0, 2q) (Shoup);
            // w[] holds 2^32 - w, so nt = mulhi * q + x * (2^32 - w) is one
            // multiply-add, and the two outputs are one subtract and one
            // three-way add
            const uint32_t x = v[k | kb];
            const uint32_t nt = (uint32_t)mad_wide(mulhi32_hd(x, ws[ti]), q, mul_wide(x, w[ti]));
            v[k] = u - nt;                 // u + t
            v[k | kb] = u + nt + q2;       // u - t + 2q
        }
    }
}

// last forward pass: [0, 4q) -> [0, q)
__host__ __device__ __forceinline__ void ntt_fwd_finish(uint32_t v[16], uint32_t q) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t x = umin_hd(v[k], v[k] - 2 * q);
        v[k] = umin_hd(x, x - q);
    }
}

// inverse pass (stages 4P+3 .. 4P, Gentleman-Sande): inputs and outputs in
// [0, 2q)
__host__ __device__ __forceinline__ void ntt_inv_pass(uint32_t v[16], const uint32_t w[15], const uint32_t ws[15],
                                                      uint32_t q) {
    const uint32_t q2 = 2 * q;
#pragma unroll
    for (int st = 3; st >= 0; --st) {
        const int kb = 8 >> st;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (k & kb) continue;
            const int ti = (1 << st) - 1 + (k >> (4 - st));
            const uint32_t u = v[k], x = v[k | kb];
            const uint32_t a = u + x;
            v[k] = umin_hd(a, a - q2);
            v[k | kb] = mul_shoup_lazy(u - x + q2, w[ti], ws[ti], q);
        }
    }
}

// last inverse pass: scale by N^-1 and reduce [0, 2q) -> [0, q)
__host__ __device__ __forceinline__ void ntt_inv_finish(uint32_t v[16], uint32_t ninv, uint32_t ninv_sh, uint32_t q) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t x = mul_shoup_lazy(v[k], ninv, ninv_sh, q);
        v[k] = umin_hd(x, x - q);
    }
}

// The twiddles of a pass on the device.  Pass 0's are the same for every
// thread: read through the constant address space they become scalar loads,
// counted in lgkmcnt rather than vmcnt, so they never wait behind LDS-DMA
// copies or other vector loads in flight.
template <int PASS>
__device__ __forceinline__ void load_pass_twiddles_dev(const uint2* tw2, int tid, uint32_t w[15], uint32_t ws[15]) {
    if constexpr (PASS == 0) {
        typedef const __attribute__((address_space(4))) uint32_t cu32;
        cu32* t = (cu32*)tw2;
#pragma unroll
        for (int k = 0; k < 15; ++k) {   // stage s at entries 2^s .. 2^(s+1) - 1
            w[k] = t[2 * (1 + k)];
            ws[k] = t[2 * (1 + k) + 1];
        }
    } else {
        load_pass_twiddles(tw2, PASS, tid, w, ws);
    }
}

// one pass with the pass index known at compile time (constant LDS offsets)
template <int PASS, bool INV>
__device__ __forceinline__ void ntt_pass_lds(uint32_t* s, const PrimeTab& P, int tid) {
    uint32_t v[16], w[15], ws[15];
    load_pass_twiddles_dev<PASS>(INV ? P.iv2 : P.fw2, tid, w, ws);
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = s[lds_pad(pass_elem(PASS, tid, k))];
    if (INV) {
        ntt_inv_pass(v, w, ws, P.q);
        if (PASS == 0) ntt_inv_finish(v, P.ninv, P.ninv_sh, P.q);
    } else {
        ntt_fwd_pass(v, w, ws, P.q);
        if (PASS == 2) ntt_fwd_finish(v, P.q);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) s[lds_pad(pass_elem(PASS, tid, k))] = v[k];
}

// whole transforms on a padded LDS buffer s[kPadN] by 256 threads (tid 0..255;
// a larger workgroup may run one transform per 256-thread group, every
// thread reaching the same barriers); the caller synchronises before (data in
// s) and after (result in s)
__device__ __forceinline__ void ntt_fwd_r16(uint32_t* s, const PrimeTab& P, int tid) {
    ntt_pass_lds<0, false>(s, P, tid);
    __syncthreads();
    ntt_pass_lds<1, false>(s, P, tid);
    __syncthreads();
    ntt_pass_lds<2, false>(s, P, tid);
    __syncthreads();
}
__device__ __forceinline__ void ntt_fwd_r16(uint32_t* s, const PrimeTab& P) { ntt_fwd_r16(s, P, (int)threadIdx.x); }

__device__ __forceinline__ void ntt_inv_r16(uint32_t* s, const PrimeTab& P, int tid) {
    ntt_pass_lds<2, true>(s, P, tid);
    __syncthreads();
    ntt_pass_lds<1, true>(s, P, tid);
    __syncthreads();
    ntt_pass_lds<0, true>(s, P, tid);
    __syncthreads();
}
__device__ __forceinline__ void ntt_inv_r16(uint32_t* s, const PrimeTab& P) { ntt_inv_r16(s, P, (int)threadIdx.x); }

// two independent transforms (primes PA, PB may differ) interleaved pass by
// pass: twice the independent work between barriers
template <int PASS, bool INV, bool SAME = false>
__device__ __forceinline__ void ntt_pass_lds2(uint32_t* sa, const PrimeTab& PA, uint32_t* sb, const PrimeTab& PB,
                                              int tid) {
    uint32_t va[16], vb[16], wa[15], wsa[15], wb[15], wsb[15];
    load_pass_twiddles_dev<PASS>(INV ? PA.iv2 : PA.fw2, tid, wa, wsa);
    if (SAME) {     // one prime for both: share the twiddle registers
#pragma unroll
        for (int t = 0; t < 15; ++t) { wb[t] = wa[t]; wsb[t] = wsa[t]; }
    } else {
        load_pass_twiddles_dev<PASS>(INV ? PB.iv2 : PB.fw2, tid, wb, wsb);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        va[k] = sa[lds_pad(pass_elem(PASS, tid, k))];
        vb[k] = sb[lds_pad(pass_elem(PASS, tid, k))];
    }
    if (INV) {
        ntt_inv_pass(va, wa, wsa, PA.q);
        ntt_inv_pass(vb, wb, wsb, PB.q);
        if (PASS == 0) {
            ntt_inv_finish(va, PA.ninv, PA.ninv_sh, PA.q);
            ntt_inv_finish(vb, PB.ninv, PB.ninv_sh, PB.q);
        }
    } else {
        ntt_fwd_pass(va, wa, wsa, PA.q);
        ntt_fwd_pass(vb, wb, wsb, PB.q);
        if (PASS == 2) {
            ntt_fwd_finish(va, PA.q);
            ntt_fwd_finish(vb, PB.q);
        }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        sa[lds_pad(pass_elem(PASS, tid, k))] = va[k];
        sb[lds_pad(pass_elem(PASS, tid, k))] = vb[k];
    }
}

// ntt_pass_lds with the pass's twiddles already loaded
template <int PASS, bool INV>
__device__ __forceinline__ void ntt_pass_lds_tw(uint32_t* s, const PrimeTab& P, int tid, const uint32_t (&w)[15],
                                                const uint32_t (&ws)[15]) {
    uint32_t v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = s[lds_pad(pass_elem(PASS, tid, k))];
    if (INV) {
        ntt_inv_pass(v, w, ws, P.q);
        if (PASS == 0) ntt_inv_finish(v, P.ninv, P.ninv_sh, P.q);
    } else {
        ntt_fwd_pass(v, w, ws, P.q);
        if (PASS == 2) ntt_fwd_finish(v, P.q);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) s[lds_pad(pass_elem(PASS, tid, k))] = v[k];
}

// ntt_pass_lds2 for two transforms in one prime with the pass's twiddles
// already loaded (w, ws as load_pass_twiddles gives them)
template <int PASS, bool INV>
__device__ __forceinline__ void ntt_pass_lds2_tw(uint32_t* sa, uint32_t* sb, const PrimeTab& P, int tid,
                                                 const uint32_t (&w)[15], const uint32_t (&ws)[15]) {
    uint32_t va[16], vb[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        va[k] = sa[lds_pad(pass_elem(PASS, tid, k))];
        vb[k] = sb[lds_pad(pass_elem(PASS, tid, k))];
    }
    if (INV) {
        ntt_inv_pass(va, w, ws, P.q);
        ntt_inv_pass(vb, w, ws, P.q);
        if (PASS == 0) {
            ntt_inv_finish(va, P.ninv, P.ninv_sh, P.q);
            ntt_inv_finish(vb, P.ninv, P.ninv_sh, P.q);
        }
    } else {
        ntt_fwd_pass(va, w, ws, P.q);
        ntt_fwd_pass(vb, w, ws, P.q);
        if (PASS == 2) {
            ntt_fwd_finish(va, P.q);
            ntt_fwd_finish(vb, P.q);
        }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        sa[lds_pad(pass_elem(PASS, tid, k))] = va[k];
        sb[lds_pad(pass_elem(PASS, tid, k))] = vb[k];
    }
}

__device__ __forceinline__ void ntt_fwd_r16x2(uint32_t* sa, const PrimeTab& PA, uint32_t* sb, const PrimeTab& PB, int tid) {
    ntt_pass_lds2<0, false>(sa, PA, sb, PB, tid);
    __syncthreads();
    ntt_pass_lds2<1, false>(sa, PA, sb, PB, tid);
    __syncthreads();
    ntt_pass_lds2<2, false>(sa, PA, sb, PB, tid);
    __syncthreads();
}
__device__ __forceinline__ void ntt_fwd_r16x2(uint32_t* sa, const PrimeTab& PA, uint32_t* sb, const PrimeTab& PB) { ntt_fwd_r16x2(sa, PA, sb, PB, (int)threadIdx.x); }

// two forward transforms in the same prime
__device__ __forceinline__ void ntt_fwd_r16x2_same(uint32_t* sa, uint32_t* sb, const PrimeTab& P, int tid) {
    ntt_pass_lds2<0, false, true>(sa, P, sb, P, tid);
    __syncthreads();
    ntt_pass_lds2<1, false, true>(sa, P, sb, P, tid);
    __syncthreads();
    ntt_pass_lds2<2, false, true>(sa, P, sb, P, tid);
    __syncthreads();
}

__device__ __forceinline__ void ntt_inv_r16x2(uint32_t* sa, const PrimeTab& PA, uint32_t* sb, const PrimeTab& PB, int tid) {
    ntt_pass_lds2<2, true>(sa, PA, sb, PB, tid);
    __syncthreads();
    ntt_pass_lds2<1, true>(sa, PA, sb, PB, tid);
    __syncthreads();
    ntt_pass_lds2<0, true>(sa, PA, sb, PB, tid);
    __syncthreads();
}
__device__ __forceinline__ void ntt_inv_r16x2(uint32_t* sa, const PrimeTab& PA, uint32_t* sb, const PrimeTab& PB) { ntt_inv_r16x2(sa, PA, sb, PB, (int)threadIdx.x); }

// 12-bit bit reversal
__device__ __forceinline__ uint32_t brev12(uint32_t x) { return __builtin_bitreverse32(x) >> (32 - kLogN); }

// NTT slot s holds the evaluation at psi^(2 brev(s) + 1) (the tables
// T.slot_exp / T.exp_slot hold the same map)
__device__ __forceinline__ uint32_t slot_exponent(int s) { return 2 * brev12((uint32_t)s) + 1; }

// slot permutation realising the automorphism X -> X^g in the NTT domain:
// NTT(tau_g(a))[s] = NTT(a)[perm_g(s)], computed in registers
__device__ __forceinline__ int perm_slot(int s, uint32_t g, const Tabs&) {
    const uint32_t e = (slot_exponent(s) * g) & (2 * kN - 1);
    return (int)brev12(e >> 1);
}

}  // namespace kspir
