// kspir_offline.h — giant-step half of the offline key derivation
// (genAutoKeyFromOffline, twosteps.cpp:1814-1863) on the GPU, for many clients
// at once.
//
// A key of AutoKeyBSGSRNS(N, crtMod, auxMod, ...) is held as u32 planes in the
// three primes q1, q2 (its crtMod rows) and auxMod (its auxMod rows):
//   state [3 rows j][2 a|b][3 primes][N].
// One derivation step (evalAutoRNSKey, twosteps.cpp:1588-1715, auxMod branch)
// maps the key at 5^(k N1) to the key at 5^((k+1) N1):
//   k_okd_digits  automorphism by 5^N1 as a gather, CRT composition mod
//                 crtMod*auxMod, 5 signed digits of 14 bits after a 14-bit
//                 rounding shift (decompose_bsgs_aux, utils.cpp:1191-1256);
//   k_okd_ks      per (client, row, prime): forward transforms of the digits
//                 and of tau(b), the product with the offline key
//                 (Montgomery), inverse transforms of a and b.
// The steps are sequential (each key is the previous one's automorphism), the
// clients and rows are independent, so a launch covers every client.
// k_okd_modswitch then switches every key to crtMod, auxMod^-1 * (x - x_aux),
// and transforms it (NTT form mod crtMod).
#pragma once
#include "kspir_device.h"

namespace kspir {

struct OkdConst {
    uint32_t auxinv[2], auxinv_sh[2];   // auxMod^-1 mod q1, q2
    uint32_t qinv_aux, qinv_aux_sh;     // crtMod^-1 mod auxMod
    uint32_t ginv;                      // (5^N1)^-1 mod 2N
};

constexpr int kOkdEll = 5;        // digits of the offline key (ell 5, PP = Bg = 2^14)
constexpr int kOkdRows = 3;       // rows of the derived key (ell 3)

// value in [0, crtMod*auxMod) congruent to xq mod crtMod and xa mod auxMod
__device__ __forceinline__ unsigned __int128 crt84(uint64_t xq, uint32_t xa, const OkdConst& K) {
    const uint32_t xq_a = (uint32_t)(xq % kAux);
    const uint32_t t = mul_shoup(sub_mod(xa, xq_a, kAux), K.qinv_aux, K.qinv_aux_sh, kAux);
    return (unsigned __int128)xq + (unsigned __int128)kQ * t;
}

// auto key words ([6][2][N] u64 per client, NTT form) -> u32 planes: the
// residues as stored (state of key 5^N1) and their inverse transforms (the
// first step's input).  grid (18 planes = row * 6 + ab * 3 + prime, nc)
__global__ void __launch_bounds__(kNttThreads) k_okd_load(const uint64_t* __restrict__ auto2, uint32_t* __restrict__ st0,
                                                         size_t st_stride, uint32_t* __restrict__ c0, Tabs T) {
    __shared__ uint32_t s[kPadN];
    const int pl = blockIdx.x, c = blockIdx.y;
    const int j = pl / 6, ab = (pl / 3) % 2, p = pl % 3;
    const uint32_t q = p == 2 ? kAux : T.p[p].q;
    const uint64_t* src = auto2 + ((size_t)c * 6 + (p == 2 ? j + kOkdRows : j)) * 2 * kN + (size_t)ab * kN;
    uint32_t* raw = st0 + (size_t)c * st_stride + (size_t)pl * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        const uint32_t v = (uint32_t)(src[x] % q);
        raw[x] = v;
        s[lds_pad(x)] = v;
    }
    __syncthreads();
    ntt_inv_r16(s, p == 2 ? T.p[4] : T.p[p]);
    uint32_t* dst = c0 + ((size_t)c * 18 + pl) * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) dst[x] = s[lds_pad(x)];
}

// offline key words ([10][2][N] u64 per client, NTT form) -> Montgomery planes
// [5 digits][2 a|b][3 primes][N]; grid (N/256, 30, nc)
__global__ void __launch_bounds__(256) k_okd_keys(const uint64_t* __restrict__ off, uint32_t* __restrict__ km,
                                                  OkdConst K) {
    const int x = blockIdx.x * 256 + threadIdx.x, pl = blockIdx.y, c = blockIdx.z;
    const int d = pl / 6, ab = (pl / 3) % 2, p = pl % 3;
    const uint32_t q = p == 0 ? kQ1 : p == 1 ? kQ2 : kAux;
    const uint64_t w = off[(((size_t)c * 10 + (p == 2 ? d + kOkdEll : d)) * 2 + ab) * kN + x] % q;
    // w * 2^32 mod q
    km[((size_t)c * 30 + pl) * kN + x] = (uint32_t)(((unsigned __int128)w << 32) % q);
}

// one derivation step, part 1: for every client and row, the automorphism by
// 5^N1 (twosteps.cpp:1622-1647, a negated word written as q - x) gathered at
// output position o from source (o * ginv) mod 2N, its composition mod
// crtMod*auxMod and the signed digits (decompose_bsgs_aux).  in: state
// [3][2][3][N] per client (coefficient form); dig [3 rows][5][3][N],
// tb [3 rows][3][N].  grid (N/256, 3 rows, nc)
__global__ void __launch_bounds__(256) k_okd_digits(const uint32_t* __restrict__ in, size_t in_stride,
                                                    uint32_t* __restrict__ dig, uint32_t* __restrict__ tb,
                                                    OkdConst K, Tabs T) {
    const int o = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, c = blockIdx.z;
    const uint32_t src = (uint32_t)(o * K.ginv) & (2 * kN - 1);
    const bool neg = src >= (uint32_t)kN;
    const int i = (int)(src & (kN - 1));
    const uint32_t* a = in + (size_t)c * in_stride + (size_t)j * 6 * kN;   // [ab][prime][N]
    const uint32_t* b = a + 3 * kN;
    const uint32_t qs[3] = {kQ1, kQ2, kAux};
    uint32_t av[3], bv[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        av[p] = a[(size_t)p * kN + i];
        bv[p] = b[(size_t)p * kN + i];
        if (neg) {
            av[p] = neg_mod(av[p], qs[p]);
            bv[p] = neg_mod(bv[p], qs[p]);
        }
    }
    const unsigned __int128 M = (unsigned __int128)kQ * kAux;
    // offset = (sum_i 2^(14 + 14 i)) * 2^13 + 2^13 for i < 5
    unsigned __int128 offset = 0;
#pragma unroll
    for (int t = 0; t < kOkdEll; ++t) offset += (unsigned __int128)1 << (14 + 14 * t);
    offset = (offset << 13) + (1u << 13);
    unsigned __int128 d = crt84(crt_q(av[0], av[1], T), av[2], K) + offset;
    if (d > M) d -= M;
    d >>= 14;
    uint32_t* dg = dig + ((size_t)c * kOkdRows + j) * kOkdEll * 3 * kN;
#pragma unroll
    for (int t = 0; t < kOkdEll; ++t) {
        const int32_t r = (int32_t)((uint32_t)(d >> (14 * t)) & 0x3FFFu) - (1 << 13);
#pragma unroll
        for (int p = 0; p < 3; ++p) dg[((size_t)t * 3 + p) * kN + o] = digit_residue(r, qs[p]);
    }
    uint32_t* tbo = tb + ((size_t)c * kOkdRows + j) * 3 * kN;
#pragma unroll
    for (int p = 0; p < 3; ++p) tbo[(size_t)p * kN + o] = bv[p];
}

// one derivation step, part 2, per (prime, row, client): (0, tau(b)) minus the
// digits times the offline key, in the NTT domain, back to coefficient form.
// out: state [3][2][3][N] per client.  grid (3 primes, 3 rows, nc)
__global__ void __launch_bounds__(kNttThreads) k_okd_ks(const uint32_t* __restrict__ dig, const uint32_t* __restrict__ tb,
                                                       const uint32_t* __restrict__ km, uint32_t* __restrict__ out,
                                                       size_t out_stride, Tabs T) {
    __shared__ uint32_t sa[kPadN];
    __shared__ uint32_t sb[kPadN];
    const int p = blockIdx.x, j = blockIdx.y, c = blockIdx.z;
    const int tid = threadIdx.x;
    const PrimeTab& P = p == 2 ? T.p[4] : T.p[p];
    const uint32_t q = P.q;
    const uint32_t* dg = dig + ((size_t)c * kOkdRows + j) * kOkdEll * 3 * kN;
    const uint32_t* kc = km + (size_t)c * 30 * kN;
    uint64_t acc_a[16], acc_b[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc_a[k] = acc_b[k] = 0;
    for (int t = 0; t < kOkdEll; ++t) {
        const uint32_t* src = dg + ((size_t)t * 3 + p) * kN;
#pragma unroll
        for (int k = 0; k < 16; ++k) sa[lds_pad(tid + 256 * k)] = src[tid + 256 * k];
        __syncthreads();
        ntt_fwd_r16(sa, P, tid);
        const uint32_t* ka = kc + ((size_t)t * 6 + p) * kN;        // [t][a][p]
        const uint32_t* kb = kc + ((size_t)t * 6 + 3 + p) * kN;    // [t][b][p]
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int x = tid + 256 * k;
            const uint64_t v = sa[lds_pad(x)];
            acc_a[k] += v * ka[x];
            acc_b[k] += v * kb[x];
        }
        __syncthreads();
    }
    const uint32_t* bsrc = tb + ((size_t)c * kOkdRows + j) * 3 * kN + (size_t)p * kN;
#pragma unroll
    for (int k = 0; k < 16; ++k) sb[lds_pad(tid + 256 * k)] = bsrc[tid + 256 * k];
    __syncthreads();
    ntt_fwd_r16(sb, P, tid);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int x = lds_pad(tid + 256 * k);
        sa[x] = neg_mod(mont_reduce(acc_a[k], q, P.mont_qinv), q);
        sb[x] = sub_mod(sb[x], mont_reduce(acc_b[k], q, P.mont_qinv), q);
    }
    __syncthreads();
    ntt_inv_r16(sa, P, tid);
    ntt_inv_r16(sb, P, tid);
    uint32_t* oa = out + (size_t)c * out_stride + ((size_t)j * 6 + p) * kN;
    uint32_t* ob = oa + 3 * kN;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int x = tid + 256 * k;
        oa[x] = sa[lds_pad(x)];
        ob[x] = sb[lds_pad(x)];
    }
}

// twosteps.cpp:1839-1862: key k of every client switched to crtMod,
// auxMod^-1 * (x - x_aux) mod crtMod, and transformed; both primes in one
// workgroup (one per 256-thread half) so the words compose to u64.
// st: [nc][n2-1][3][2][3][N]; giant: [nc][n2-1][3][2][N] u64.
// grid (n2-1 keys, 6 = row * 2 + ab, nc), 512 threads
__global__ void __launch_bounds__(512) k_okd_modswitch(const uint32_t* __restrict__ st, size_t st_stride,
                                                       uint64_t* __restrict__ giant, int nkeys, OkdConst K, Tabs T) {
    __shared__ uint32_t s[2][kPadN];
    const int kk = blockIdx.x, jab = blockIdx.y, c = blockIdx.z;
    const int tid = threadIdx.x;
    const int half = __builtin_amdgcn_readfirstlane(tid >> 8), ht = tid & 255;
    const PrimeTab& P = T.p[half];
    const uint32_t q = P.q;
    const uint32_t* src = st + (size_t)c * st_stride + ((size_t)kk * 6 + jab) * 3 * kN;   // [prime][N]
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
        const int x = ht + 256 * k;
        uint32_t xa = src[2 * kN + x];
        if (xa >= q) xa -= q;   // auxMod < 2 q
        const uint32_t v = sub_mod(src[(size_t)half * kN + x], xa, q);
        s[half][lds_pad(x)] = mul_shoup(v, K.auxinv[half], K.auxinv_sh[half], q);
    }
    __syncthreads();
    ntt_fwd_r16(s[half], P, ht);
    uint64_t* dst = giant + (((size_t)c * nkeys + kk) * 6 + jab) * kN;
    for (int x = tid; x < kN; x += 512) dst[x] = crt_q(s[0][lds_pad(x)], s[1][lds_pad(x)], T);
}

// chain states 1..n2-2 -> u64 words [nc][n2-2][6][2][N]: rows 0..2 mod crtMod,
// rows 3..5 the auxMod residues.  grid (N/256, (n2-2) * 12, nc)
__global__ void __launch_bounds__(256) k_okd_export(const uint32_t* __restrict__ st, size_t st_stride,
                                                    uint64_t* __restrict__ chain, int nchain, Tabs T) {
    const int x = blockIdx.x * 256 + threadIdx.x, c = blockIdx.z;
    const int kk = blockIdx.y / 12, e = (blockIdx.y / 2) % 6, ab = blockIdx.y % 2;
    const int j = e % kOkdRows;
    const uint32_t* src = st + (size_t)c * st_stride + ((size_t)(kk + 1) * 6 + j * 2 + ab) * 3 * kN;
    uint64_t v;
    if (e < kOkdRows) v = crt_q(src[x], src[kN + x], T);
    else v = src[2 * kN + x];
    chain[((((size_t)c * nchain + kk) * 6 + e) * 2 + ab) * kN + x] = v;
}

}  // namespace kspir
