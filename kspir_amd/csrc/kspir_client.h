// kspir_client.h — client side on the GPU: the glibc rand() stream, the
// secret, query_bsgs_rns and the RGSW / BSGS / packing key generation of
// test-pir.cpp:20-24,68-89, for many clients at once.
//
// Randomness: each client's rand() outputs (after srand(seed)) are laid out
// as one stream in HBM, produced by jump-ahead of the generator's 31-word
// window (kspir_rng.h): k_rng_windows walks one window per chunk of
// kRngChunk outputs with a precomputed M^kRngChunk, k_rng_stream expands the
// chunks.  Every sampler of the reference then reads its draws at a fixed
// offset of the stream, in the reference's consumption order:
//   sample_random(a, q) (samples.cpp:9-18): a = ((o_k << 31) | o_{k+1}) % q
//   sample_guass (samples.cpp:56-65) and the secret (secret.cpp:13-60):
//     t = o_k % 3 - 1
// All ring arithmetic is in u32 residues (q1, q2 for crtMod, bsMod), NTTs as
// in the server kernels; words leave as u64 in the reference's layouts.
#pragma once
#include "kspir_device.h"

namespace kspir {

constexpr int kRngLag = 31;
constexpr int kRngChunk = 31 * 64;   // outputs per generator lane

// one workgroup (64 threads) per client: window of every chunk,
// win[c][ch] = M^(ch * kRngChunk) S_344(seed_c)
__global__ void __launch_bounds__(64) k_rng_windows(const uint32_t* __restrict__ seedw,
                                                    const uint32_t* __restrict__ jump, uint32_t* __restrict__ win,
                                                    int nchunks) {
    __shared__ uint32_t J[kRngLag * kRngLag];
    __shared__ uint32_t w[32];
    const int c = blockIdx.x, t = threadIdx.x;
    for (int i = t; i < kRngLag * kRngLag; i += 64) J[i] = jump[i];
    if (t < kRngLag) w[t] = seedw[c * kRngLag + t];
    __syncthreads();
    uint32_t* out = win + (size_t)c * nchunks * kRngLag;
    for (int ch = 0; ch < nchunks; ++ch) {
        uint32_t nv = 0;
        if (t < kRngLag) {
            out[(size_t)ch * kRngLag + t] = w[t];
#pragma unroll 31
            for (int j = 0; j < kRngLag; ++j) nv += J[t * kRngLag + j] * w[j];
        }
        __syncthreads();
        if (t < kRngLag) w[t] = nv;
        __syncthreads();
    }
}

// one lane per (client, chunk): kRngChunk outputs, 31 per step of the
// in-register window (r[t] = r[t-31] + r[t-3] updated in place), staged in
// LDS so each store writes one lane's 31 consecutive outputs (124 bytes).
// stream: [nc][nchunks * kRngChunk] u32.  1-D grid of 64-thread groups.
__global__ void __launch_bounds__(64) k_rng_stream(const uint32_t* __restrict__ win, uint32_t* __restrict__ stream,
                                                   int nc, int nchunks) {
    __shared__ uint32_t T[64][33];     // row pad: lane-row writes hit distinct banks
    const int lane = threadIdx.x;
    const int g = blockIdx.x * 64 + lane;
    const int total = nc * nchunks;
    const bool live = g < total;
    uint32_t R[kRngLag];
#pragma unroll
    for (int i = 0; i < kRngLag; ++i) R[i] = live ? win[(size_t)g * kRngLag + i] : 0u;
    const int g0 = blockIdx.x * 64;
    for (int b = 0; b < kRngChunk / kRngLag; ++b) {
#pragma unroll
        for (int j = 0; j < kRngLag; ++j) {
            R[j] += (j < 3) ? R[28 + j] : R[j - 3];
            T[lane][j] = R[j] >> 1;
        }
        __syncthreads();
        for (int l = 0; l < 64; ++l) {
            const int gl = g0 + l;
            if (gl < total && lane < kRngLag) {
                const size_t c = gl / nchunks, ch = gl % nchunks;
                stream[(c * nchunks + ch) * kRngChunk + (size_t)b * kRngLag + lane] = T[l][lane];
            }
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int32_t ternary(uint32_t o) { return (int32_t)(o % 3u) - 1; }

// ((hi << 31) | lo) % q for the two draws at o
__device__ __forceinline__ uint64_t uniform_draw(const uint32_t* o, uint64_t q) {
    return ((((uint64_t)o[0]) << 31) | (uint64_t)o[1]) % q;
}

// residue mod q of a ternary value
__device__ __forceinline__ uint32_t tern_res(int32_t t, uint32_t q) { return t < 0 ? q - 1 : (uint32_t)t; }

// Montgomery form x * 2^32 mod q from x < q, r2 = 2^64 mod q
__device__ __forceinline__ uint32_t to_mont(uint32_t x, uint32_t r2, const PrimeTab& P) {
    return mont_reduce((uint64_t)x * r2, P.q, P.mont_qinv);
}

struct ClientConsts {
    uint32_t r2[3];          // 2^64 mod q1, q2, bsMod
    uint32_t bs_mod_q[2];    // bsMod mod q1, q2
};

// secret.cpp:13-60: ternary secret from draws [0, N) of the stream.
// scoef[c][N] int32 in {-1, 0, 1}; sm[c][3][N] NTT of the secret mod q1, q2
// and (lifted) mod bsMod, Montgomery form.  grid nc, 256 threads.
__global__ void __launch_bounds__(kNttThreads) k_client_secret(const uint32_t* __restrict__ stream, size_t stride,
                                                               int32_t* __restrict__ scoef, uint32_t* __restrict__ sm,
                                                               ClientConsts K, Tabs T) {
    __shared__ uint32_t s[3][kPadN];
    const int c = blockIdx.x;
    const uint32_t* o = stream + (size_t)c * stride;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        const int32_t t = ternary(o[x]);
        scoef[(size_t)c * kN + x] = t;
        for (int p = 0; p < 3; ++p) s[p][lds_pad(x)] = tern_res(t, T.p[p].q);
    }
    __syncthreads();
    ntt_fwd_r16x2(s[0], T.p[0], s[1], T.p[1]);
    ntt_fwd_r16(s[2], T.p[2]);
    for (int x = threadIdx.x; x < kN; x += kNttThreads)
        for (int p = 0; p < 3; ++p) sm[((size_t)c * 3 + p) * kN + x] = to_mont(s[p][lds_pad(x)], K.r2[p], T.p[p]);
}

// the same from given secret coefficients (ternary, as int32): grid nc, 256 threads
__global__ void __launch_bounds__(kNttThreads) k_client_secret_given(const int32_t* __restrict__ scoef,
                                                                     uint32_t* __restrict__ sm, ClientConsts K, Tabs T) {
    __shared__ uint32_t s[3][kPadN];
    const int c = blockIdx.x;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        const int32_t t = scoef[(size_t)c * kN + x];
        for (int p = 0; p < 3; ++p) s[p][lds_pad(x)] = tern_res(t, T.p[p].q);
    }
    __syncthreads();
    ntt_fwd_r16x2(s[0], T.p[0], s[1], T.p[1]);
    ntt_fwd_r16(s[2], T.p[2]);
    for (int x = threadIdx.x; x < kN; x += kNttThreads)
        for (int p = 0; p < 3; ++p) sm[((size_t)c * 3 + p) * kN + x] = to_mont(s[p][lds_pad(x)], K.r2[p], T.p[p]);
}

// automorphic(s, index) of a ternary secret at coefficient x (utils.cpp:1355-1379):
// coefficient i goes to i*index mod 2N, negated past N
__device__ __forceinline__ int32_t auto_secret(const int32_t* sc, int x, uint32_t index_inv) {
    const uint32_t it = ((uint32_t)x * index_inv) & (2 * kN - 1);
    return it < (uint32_t)kN ? sc[it] : -sc[it - kN];
}

enum ClientJobType : int32_t {
    kJobRlwe = 0,      // encrypt_rlwe / encrypt_special_rlwe of msg = mf * m(x) (+ num * err)
    kJobRgswNeg = 1,   // encrypt_rlwe of -(NTT(mf * m) * s)
    kJobBaby = 2,      // encrypt_rns_bsgs_autokey of (mf * tau(s) mod Q, mf * tau(s) mod bsMod)
};

struct ClientJob {
    int32_t type;
    int32_t client;
    uint32_t draw;       // stream offset of this encryption's first draw
    int32_t msg;         // >= 1: m = automorphic(s, msg) (msg_inv its inverse mod 2N); <= 0: m = X^(-(-msg))
    uint32_t msg_inv;
    int32_t msg_plus;    // 1: for msg <= 0, m = X^(-msg) instead (RGSW keyGen without is_reverse)
    uint32_t mf[2];      // gadget factor mod q1, q2
    uint32_t mf_bs;      // gadget factor mod bsMod (baby keys)
    uint32_t num;        // error multiplier (encrypt_special_rlwe), 1 otherwise
    uint64_t* a;         // outputs, [N] u64 words: a, b mod crtMod; a2, b2 mod bsMod
    uint64_t* b;
    uint64_t* a2;
    uint64_t* b2;
};

// message coefficient m(x) in {-1, 0, 1}: tau_index(s) or X^(-idx) (lwe.cpp:297-352:
// idx 0 -> 1, else coefficient N - idx = -1)
__device__ __forceinline__ int32_t job_msg(const ClientJob& J, const int32_t* sc, int x) {
    if (J.msg >= 1) return auto_secret(sc, x, J.msg_inv);
    const int idx = -J.msg;
    if (J.msg_plus) return x == idx ? 1 : 0;
    if (idx == 0) return x == 0 ? 1 : 0;
    return x == kN - idx ? -1 : 0;
}

// signed coefficient times a residue factor, mod q
__device__ __forceinline__ uint32_t smul(int32_t m, uint32_t f, uint32_t q) {
    return m == 0 ? 0u : (m > 0 ? f : (f ? q - f : 0u));
}

// One RLWE encryption per workgroup (encrypt.cpp:53-104, 391-423, 300-387,
// lwe.cpp:297-352, 472-501, 1025-1101).  a (sample_random mod Q) is taken as
// NTT form; b = a*s + NTT(num*err + msg) (linear in the NTT), or for the
// negated RGSW rows b = a*s + NTT(err) - NTT(msg)*s; the baby keys add the
// bsMod half a2, b2 = a2*s_bs + NTT_bs(lift(err) + msg_bs).  grid njobs, 256 threads.
__global__ void __launch_bounds__(kNttThreads) k_client_rlwe(const ClientJob* __restrict__ jobs,
        const uint32_t* __restrict__ stream, size_t stride, const int32_t* __restrict__ scoef,
        const uint32_t* __restrict__ sm, ClientConsts K, Tabs T) {
    __shared__ uint32_t s[4][kPadN];
    const ClientJob J = jobs[blockIdx.x];
    const uint32_t* o = stream + (size_t)J.client * stride + J.draw;
    const int32_t* sc = scoef + (size_t)J.client * kN;
    const uint32_t* smc = sm + (size_t)J.client * 3 * kN;
    const int tid = threadIdx.x;
    // err (and the bsMod copy for the baby keys)
    const uint32_t* oe = o + (J.type == kJobBaby ? 4 * kN : 2 * kN);
    for (int x = tid; x < kN; x += kNttThreads) {
        const int32_t e = ternary(oe[x]);
        const int32_t m = job_msg(J, sc, x);
        for (int p = 0; p < 2; ++p) {
            const uint32_t q = T.p[p].q;
            const uint32_t ev = tern_res(e, q);
            const uint32_t en = (uint32_t)(((uint64_t)ev * J.num) % q);
            const uint32_t mv = smul(m, J.mf[p], q);
            if (J.type == kJobRgswNeg) {
                s[p][lds_pad(x)] = ev;
                s[2 + p][lds_pad(x)] = mv;
            } else {
                s[p][lds_pad(x)] = add_mod(en, mv, q);
            }
        }
        if (J.type == kJobBaby) s[2][lds_pad(x)] = add_mod(tern_res(e, kBs), smul(m, J.mf_bs, kBs), kBs);
    }
    __syncthreads();
    ntt_fwd_r16x2(s[0], T.p[0], s[1], T.p[1]);
    if (J.type == kJobRgswNeg) ntt_fwd_r16x2(s[2], T.p[0], s[3], T.p[1]);
    if (J.type == kJobBaby) ntt_fwd_r16(s[2], T.p[2]);
    for (int x = tid; x < kN; x += kNttThreads) {
        const uint64_t A = uniform_draw(o + 2 * x, kQ);
        uint32_t bq[2];
        for (int p = 0; p < 2; ++p) {
            const PrimeTab& P = T.p[p];
            const uint32_t ap = (uint32_t)(A % P.q);
            const uint32_t sx = smc[(size_t)p * kN + x];
            uint32_t v = add_mod(mont_reduce((uint64_t)ap * sx, P.q, P.mont_qinv), s[p][lds_pad(x)], P.q);
            if (J.type == kJobRgswNeg)
                v = sub_mod(v, mont_reduce((uint64_t)s[2 + p][lds_pad(x)] * sx, P.q, P.mont_qinv), P.q);
            bq[p] = v;
        }
        J.a[x] = A;
        J.b[x] = crt_q(bq[0], bq[1], T);
        if (J.type == kJobBaby) {
            const uint64_t A2 = uniform_draw(o + 2 * kN + 2 * x, kBs);
            const PrimeTab& P = T.p[2];
            const uint32_t v = add_mod(mont_reduce(A2 * smc[(size_t)2 * kN + x], P.q, P.mont_qinv), s[2][lds_pad(x)],
                                       P.q);
            J.a2[x] = A2;
            J.b2[x] = v;
        }
    }
}

// query_bsgs_rns (twosteps.cpp:279-318) with encrypt_rns_bsgs
// (encrypt.cpp:443-495, 118-152): message slots e0, e1 set to 1 in the
// plaintext NTT domain, INTT mod p, each coefficient encoded by enc_tab
// (round(m * Q / p) in the reference's long double arithmetic); then
// c0 = (a0, a0*s + NTT(m*bsMod + err)) mod Q, c1 = (a1, a1*s + NTT(lift(err)))
// mod bsMod, all four returned to coefficient form.  Draws, from `base` of
// the client's stream (N after the secret): a0 [0, 2N), a1 [2N, 4N),
// err [4N, 5N).  out: [nc][4][N] (c0.a, c0.b, c1.a, c1.b).  grid nc, 256 threads.
__global__ void __launch_bounds__(kNttThreads) k_client_query(const uint32_t* __restrict__ stream, size_t stride,
        size_t base, const uint32_t* __restrict__ sm, const int32_t* __restrict__ slots,
        const uint64_t* __restrict__ enc_tab, uint64_t* __restrict__ out, ClientConsts K, Tabs T) {
    __shared__ uint32_t s[4][kPadN];
    const int c = blockIdx.x, tid = threadIdx.x;
    const uint32_t* o = stream + (size_t)c * stride + base;   // a0 at o[0], a1 at o[2N], err at o[4N]
    const uint32_t* smc = sm + (size_t)c * 3 * kN;
    uint64_t* oc = out + (size_t)c * 4 * kN;
    const int e0 = slots[2 * c], e1 = slots[2 * c + 1];
    for (int x = tid; x < kN; x += kNttThreads) s[3][lds_pad(x)] = (x == e0 || x == e1) ? 1u : 0u;
    __syncthreads();
    ntt_inv_r16(s[3], T.p[3]);
    // NTT_Q(m * bsMod + err) in planes 0, 1; NTT_bs(lift(err)) in plane 2
    for (int x = tid; x < kN; x += kNttThreads) {
        const uint64_t M = enc_tab[s[3][lds_pad(x)]];
        const int32_t e = ternary(o[4 * kN + x]);
        for (int p = 0; p < 2; ++p) {
            const uint32_t q = T.p[p].q;
            const uint32_t mv = (uint32_t)(((M % q) * K.bs_mod_q[p]) % q);
            s[p][lds_pad(x)] = add_mod(mv, tern_res(e, q), q);
        }
        s[2][lds_pad(x)] = tern_res(e, kBs);
    }
    __syncthreads();
    ntt_fwd_r16x2(s[0], T.p[0], s[1], T.p[1]);
    ntt_fwd_r16(s[2], T.p[2]);
    // b parts in the NTT domain, in place; a parts kept for their INTTs below
    for (int x = tid; x < kN; x += kNttThreads) {
        const uint64_t A0 = uniform_draw(o + 2 * x, kQ);
        for (int p = 0; p < 2; ++p) {
            const PrimeTab& P = T.p[p];
            s[p][lds_pad(x)] = add_mod(mont_reduce((uint64_t)(uint32_t)(A0 % P.q) * smc[(size_t)p * kN + x], P.q,
                                                   P.mont_qinv), s[p][lds_pad(x)], P.q);
        }
        const uint64_t A1 = uniform_draw(o + 2 * kN + 2 * x, kBs);
        const PrimeTab& P = T.p[2];
        s[2][lds_pad(x)] = add_mod(mont_reduce(A1 * smc[(size_t)2 * kN + x], P.q, P.mont_qinv), s[2][lds_pad(x)], P.q);
    }
    __syncthreads();
    ntt_inv_r16x2(s[0], T.p[0], s[1], T.p[1]);
    ntt_inv_r16(s[2], T.p[2]);
    for (int x = tid; x < kN; x += kNttThreads) {
        oc[1 * kN + x] = crt_q(s[0][lds_pad(x)], s[1][lds_pad(x)], T);
        oc[3 * kN + x] = s[2][lds_pad(x)];
    }
    __syncthreads();
    // the a parts
    for (int x = tid; x < kN; x += kNttThreads) {
        const uint64_t A0 = uniform_draw(o + 2 * x, kQ);
        s[0][lds_pad(x)] = (uint32_t)(A0 % kQ1);
        s[1][lds_pad(x)] = (uint32_t)(A0 % kQ2);
        s[2][lds_pad(x)] = (uint32_t)uniform_draw(o + 2 * kN + 2 * x, kBs);
    }
    __syncthreads();
    ntt_inv_r16x2(s[0], T.p[0], s[1], T.p[1]);
    ntt_inv_r16(s[2], T.p[2]);
    for (int x = tid; x < kN; x += kNttThreads) {
        oc[0 * kN + x] = crt_q(s[0][lds_pad(x)], s[1][lds_pad(x)], T);
        oc[2 * kN + x] = s[2][lds_pad(x)];
    }
}

// secret coefficients mod crtMod as the reference stores them: [nc][N] u64
__global__ void __launch_bounds__(256) k_client_secret_words(const int32_t* __restrict__ scoef, uint64_t* __restrict__ out,
                                                             size_t total) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int32_t t = scoef[i];
        out[i] = t < 0 ? kQ - 1 : (uint64_t)t;
    }
}

}  // namespace kspir
