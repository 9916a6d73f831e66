// kspir_kernels.h — the HIP kernels of the KsPIR online answer (gfx950).
//
// Pipeline for one query (r database blocks, BSGS split N1 x N2, N1*N2 = N/2):
//   baby_prep      decompose the query's a (80-bit CRT, 3 digits), NTT the
//                  digits and b, modulus-switch the query (rotation 0)
//   baby_rot       one workgroup per (rotation, a|b): permuted digit x key
//                  products in 3 primes, INTT mod bsMod, NTT of the bsMod
//                  part in q1/q2, modulus switch -> packed rotation words
//   baby_mid_*     second half: decompose the un-switched rotation N1/2-1
//   db_product     out[i][z] = sum_k db[i][k][z] * ct[k][z] per residue,
//                  u64 accumulation, one lane per NTT slot
//   giant_coef / giant_ks     INTT, automorphism, gadget decomposition and
//                  key switch of the N2-1 giant-step rows of every block
//   ext_digit / ext_sum    RGSW external product
//   pack_coef / pack_digit / pack_combine   one level of the packing tree
//   response       INTT + CRT to the u64 response words
// All polynomials are u32 residue planes of N = 4096 words.
#pragma once
#include "kspir_device.h"

namespace kspir {

// ------------------------------------------------------------ batch layout

// device pointers of one client's resident keys
struct KeyPtrs {
    const uint32_t* baby;      // [N1/2][3][2][3][N] Montgomery residues
    const uint32_t* giant;     // [N2-1][2][2][2][N]
    const uint32_t* packing;   // [log2 r][4][2][2][N]
    const uint32_t* rgsw;      // [4][2][2][N]
    const uint16_t* perms;     // [N1][N]
};

// A launch that spans several databases (one segment per database, all of
// one N1): each segment owns a contiguous range of the launch's workgroups,
// wg_end holding the running totals.
constexpr int kMaxSeg = 8;

// database-pass segments: database limbs and their row range of this chunk,
// the column tiles and (sweep-relative) queries of the segment
struct DbSegs {
    int n;                        // 0: a single-database launch (the kernel's own arguments)
    int wg_end[kMaxSeg];
    const int8_t* dbl[kMaxSeg];
    int Rt[kMaxSeg], R[kMaxSeg], r0[kMaxSeg];
    int ct_lo[kMaxSeg], ct_hi[kMaxSeg], q_lo[kMaxSeg], q_hi[kMaxSeg];
};

// giant-step segments: first query, blocks of this chunk per query, and the
// external-product output buffer of the segment's queries
struct GiantSegs {
    int n;                        // 0: a single-database launch
    int wg_end[kMaxSeg];
    int q_lo[kMaxSeg], nb[kMaxSeg];
    uint32_t* out[kMaxSeg];
};

// packing-level segments: the queries of a level when databases of several
// r share the launch (a level of half-size i runs for the queries whose r is
// at least 2 i), one segment per database
struct PackSegs {
    int n;                        // 0: every query of the launch
    int wg_end[kMaxSeg];
    int q_lo[kMaxSeg];
};

// segment of workgroup job L (uniform over the workgroup); L becomes
// segment-relative
__device__ __forceinline__ int seg_of(int& L, const int* wg_end, int n) {
    int s = 0;
    while (s + 1 < n && L >= wg_end[s]) ++s;
    if (s) L -= wg_end[s - 1];
    return s;
}

// per-query strides (in 32-bit words unless noted) of the answer workspace;
// query q of a batch uses base + q * stride for every buffer
struct QS {
    size_t D, bhat, mid, icoef, ctw, dbout, gcoef, gks, fd, ecoef, epart, ext, pcoef, ppart;
    size_t qin64, resp64;   // 64-bit words
};

// ------------------------------------------------------------------ helpers

// value in [0, Q*bsMod) congruent to xq mod Q and xb mod bsMod (crt_inv_bsgs)
__device__ __forceinline__ unsigned __int128 crt80(uint64_t xq, uint32_t xb, const Tabs& T) {
    uint32_t xq_bs = (uint32_t)(xq % kBs);
    uint32_t t = mul_shoup(sub_mod(xb, xq_bs, kBs), T.qinv_bs, T.qinv_bs_sh, kBs);
    return (unsigned __int128)xq + (unsigned __int128)kQ * t;
}

// coefficient s of tau_g(a), a given by its residues in two planes; returns
// the canonical value in [0, Q).  ginv = g^-1 mod 2N.
__device__ __forceinline__ uint64_t auto_coeff(const uint32_t* a1, const uint32_t* a2, int s,
                                               uint32_t ginv, const Tabs& T) {
    uint32_t i0 = ((uint32_t)s * ginv) & (2 * kN - 1);
    bool neg = i0 >= (uint32_t)kN;
    uint32_t i = neg ? i0 - kN : i0;
    uint64_t x = crt_q(a1[i], a2[i], T);
    return neg ? (x ? kQ - x : 0) : x;
}

// the same from two padded LDS planes, with the sign of the negacyclic wrap
// applied to the residues (-x mod q_p) before the composition: the
// composition of (-x1, -x2) is Q - x, or 0 for x = 0, exactly the value the
// 64-bit negation gives, in 32-bit operations
__device__ __forceinline__ uint64_t auto_coeff_lds(const uint32_t* s1, const uint32_t* s2, int s, uint32_t ginv,
                                                   const Tabs& T) {
    const uint32_t i0 = ((uint32_t)s * ginv) & (2 * kN - 1);
    const bool neg = i0 >= (uint32_t)kN;
    const int i = (int)(i0 & (kN - 1));
    uint32_t x1 = s1[lds_pad(i)], x2 = s2[lds_pad(i)];
    if (neg) {
        x1 = neg_mod(x1, kQ1);
        x2 = neg_mod(x2, kQ2);
    }
    return crt_q(x1, x2, T);
}

// 5^e mod 2N
__device__ __forceinline__ uint32_t pow5_mod2n(uint32_t e) {
    uint32_t r = 1, b = 5;
    while (e) {
        if (e & 1) r = (r * b) & (2 * kN - 1);
        b = (b * b) & (2 * kN - 1);
        e >>= 1;
    }
    return r;
}

// X^x at NTT slot s, in Montgomery form (psi_pow holds psi^e * 2^32 mod q)
__device__ __forceinline__ uint32_t xpow_m(int s, uint32_t x, const PrimeTab& P, const Tabs& T) {
    return P.psi_pow[(slot_exponent(s) * x) & (2 * kN - 1)];
}

__device__ __forceinline__ uint32_t mont_mul(uint32_t a, uint32_t bm, const PrimeTab& P) {
    return mont_reduce((uint64_t)a * bm, P.q, P.mont_qinv);
}

// ------------------------------------------------------------ key uploads

// Q-keys: rows of reference RLWE words [row][2][N] (mod Q) -> [row][2][2 primes][N] Montgomery
__global__ void __launch_bounds__(256) k_keys_q(uint32_t* __restrict__ out, const uint64_t* __restrict__ in,
                                                int rows, Tabs T) {
    const size_t total = (size_t)rows * 2 * kN;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const size_t rab = idx / kN;
        const int s = (int)(idx % kN);
        const uint64_t v = in[idx];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const uint32_t q = T.p[p].q;
            const uint64_t m = ((v % q) << 32) % q;
            out[(rab * 2 + p) * kN + s] = (uint32_t)m;
        }
    }
}

// baby keys: [nb][6][2][N] (rows 0..2 mod Q, 3..5 mod bsMod) -> [nb][3][2][3 primes][N],
// key i stored in the slot order of its rotation: word s goes to slot
// perms[i+1][s] (perms [N1][N], each row a permutation of the slots)
__global__ void __launch_bounds__(256) k_keys_baby(uint32_t* __restrict__ out, const uint64_t* __restrict__ in,
                                                   const uint16_t* __restrict__ perms, int nb, Tabs T) {
    const size_t total = (size_t)nb * 3 * 2 * kN;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)(idx % kN);
        const int ab = (int)((idx / kN) % 2);
        const int j = (int)((idx / (2 * kN)) % 3);
        const size_t key = idx / (6 * (size_t)kN);
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const int row = p < 2 ? j : j + 3;
            const uint64_t v = in[((key * 6 + row) * 2 + ab) * kN + s];
            const uint32_t q = T.p[p].q;
            const uint64_t m = ((v % q) << 32) % q;
            out[(((key * 3 + j) * 2 + ab) * 3 + p) * kN + perms[(key + 1) * kN + s]] = (uint32_t)m;
        }
    }
}

// database words, reference layout [z][N2][N1] -> slot-minor layout [N2][N1][z]
__global__ void __launch_bounds__(256) k_db_transpose(uint64_t* __restrict__ out, const uint64_t* __restrict__ in,
                                                      int rows /* N2*N1 */) {
    __shared__ uint64_t tile[32][33];
    const int z0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int k = ty; k < 32; k += 8) tile[k][tx] = in[(size_t)(z0 + k) * rows + r0 + tx];
    __syncthreads();
    for (int k = ty; k < 32; k += 8) out[(size_t)(r0 + k) * kN + z0 + tx] = tile[tx][k];
}

// slot-minor layout [N2][N1][z] -> reference layout [z][N2][N1] (export)
__global__ void __launch_bounds__(256) k_db_untranspose(uint64_t* __restrict__ out, const uint64_t* __restrict__ in,
                                                        int rows /* N2*N1 */) {
    __shared__ uint64_t tile[32][33];
    const int z0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int k = ty; k < 32; k += 8) tile[k][tx] = in[(size_t)(r0 + k) * kN + z0 + tx];
    __syncthreads();
    for (int k = ty; k < 32; k += 8) out[(size_t)(z0 + k) * rows + r0 + tx] = tile[tx][k];
}

// Server preprocessing of one database block, one row kk = k*N1 + i per
// workgroup: database_tobsgsntt (twosteps.cpp:42-179) then database_tocrt
// (crt.cpp:859-885).  The row gathers the diagonal of the record matrix
// (result[kk][j] = data[(j - kk) mod N/2][j], mirrored for j >= N/2), then
// query_encode and permutation row i as one gather table gtab[i][x];
// INTT mod p, centre lift to Q (x > p/2 -> x + Q - p), NTT mod q1 and q2.
// pack != 0: words (x mod q1) | (x mod q2) << 32 in the slot-minor layout
// [N2*N1][N]; pack == 0: x mod Q (the result of database_tobsgsntt).
// rec: [N][N/2] record values < p (sample_database_bsgs layout).
__global__ void __launch_bounds__(kNttThreads) k_db_prep(const uint64_t* __restrict__ rec,
        const uint16_t* __restrict__ gtab, uint64_t* __restrict__ out, int n1, int pack, Tabs T) {
    __shared__ uint32_t s[2][kPadN];
    constexpr int H = kN / 2;
    constexpr int PER = kN / kNttThreads;
    const int kk = blockIdx.x, i = kk % n1, tid = threadIdx.x;
    const uint16_t* g = gtab + (size_t)i * kN;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = tid + e * kNttThreads;
        const int j = g[x];
        uint64_t v;
        if (j < H) {
            const int row = (j - kk + H) & (H - 1);
            v = rec[(size_t)row * H + j];
        } else {
            const int jj = kN - 1 - j;
            const int row = (jj - kk + H) & (H - 1);
            v = rec[(size_t)(kN - 1 - row) * H + jj];
        }
        s[0][lds_pad(x)] = (uint32_t)v;
    }
    __syncthreads();
    ntt_inv_r16(s[0], T.p[3]);
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = tid + e * kNttThreads;
        const uint32_t v = s[0][lds_pad(x)];
        const bool lift = v > kP / 2;
        s[0][lds_pad(x)] = lift ? v + kQ1 - kP : v;
        s[1][lds_pad(x)] = lift ? v + kQ2 - kP : v;
    }
    __syncthreads();
    ntt_fwd_r16x2(s[0], T.p[0], s[1], T.p[1]);
    uint64_t* o = out + (size_t)kk * kN;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = tid + e * kNttThreads;
        const uint32_t x1 = s[0][lds_pad(x)], x2 = s[1][lds_pad(x)];
        o[x] = pack ? ((uint64_t)x2 << 32 | x1) : crt_q(x1, x2, T);
    }
}

// ---------------------------------------------------------- baby steps

// rotation-0 and first-half digit preparation from the host query.
// q: [2][2][N] u64 (mod Q part a,b; mod bsMod part a,b), coefficient form.
// blocks 0..8   digit (j = b/3, prime p = b%3) of decompose_bsgs(a) -> D[j][p]
// blocks 9..11  NTT of b in prime p -> bhat[p]
// blocks 12..15 modulus-switched query (a|b, prime) -> ctw[0][ab][p][z]
__global__ void __launch_bounds__(kNttThreads) k_baby_prep(const uint64_t* __restrict__ q,
        uint32_t* __restrict__ D, uint32_t* __restrict__ bhat, uint32_t* __restrict__ ctw, int n1, QS S, Tabs T) {
    __shared__ uint32_t s[kPadN];
    const int b = blockIdx.x;
    const size_t qi = blockIdx.y;
    q += qi * S.qin64;
    D += qi * S.D;
    bhat += qi * S.bhat;
    ctw += qi * S.ctw;
    const uint64_t* qa = q;
    const uint64_t* qb = q + kN;
    const uint64_t* ba = q + 2 * kN;
    const uint64_t* bb = q + 3 * kN;
    if (b < 9) {
        const int j = b / 3, p = b % 3;
        const uint32_t qq = T.p[p].q;
        for (int x = threadIdx.x; x < kN; x += kNttThreads) {
            unsigned __int128 v = crt80(qa[x], (uint32_t)ba[x], T);
            s[lds_pad(x)] = digit_residue(bsgs_digit(v, j), qq);
        }
        __syncthreads();
        ntt_fwd_r16(s, T.p[p]);
        for (int x = threadIdx.x; x < kN; x += kNttThreads) D[(j * 3 + p) * kN + x] = s[lds_pad(x)];
    } else if (b < 12) {
        const int p = b - 9;
        const uint32_t qq = T.p[p].q;
        for (int x = threadIdx.x; x < kN; x += kNttThreads)
            s[lds_pad(x)] = p < 2 ? (uint32_t)(qb[x] % qq) : (uint32_t)bb[x];
        __syncthreads();
        ntt_fwd_r16(s, T.p[p]);
        for (int x = threadIdx.x; x < kN; x += kNttThreads) bhat[p * kN + x] = s[lds_pad(x)];
    } else {
        // twosteps.cpp:1074-1086: (c_Q - c_bs) * bsMod^-1 mod Q, then NTT
        const int ab = (b - 12) >> 1, p = (b - 12) & 1;
        const uint32_t qq = T.p[p].q;
        const uint64_t* xq = ab ? qb : qa;
        const uint64_t* xb = ab ? bb : ba;
        for (int x = threadIdx.x; x < kN; x += kNttThreads) {
            uint32_t v = sub_mod((uint32_t)(xq[x] % qq), (uint32_t)xb[x], qq);
            s[lds_pad(x)] = mul_shoup(v, T.bsinv_q[p], T.bsinv_q_sh[p], qq);
        }
        __syncthreads();
        ntt_fwd_r16(s, T.p[p]);
        for (int x = threadIdx.x; x < kN; x += kNttThreads)
            ctw[((size_t)(0 * 2 + ab) * 2 + p) * kN + x] = s[lds_pad(x)];
    }
}

// workgroup id -> linear job id; when the job count is a multiple of 8 the
// ids are dealt so that consecutive jobs run on one XCD (the hardware hands
// consecutive workgroups to the 8 XCDs in turn), where jobs of one query
// share that query's data in L2
__device__ __forceinline__ int xcd_job(int bid, int total) {
    return (total & 7) ? bid : (bid & 7) * (total >> 3) + (bid >> 3);
}

// One baby-step rotation of one query per workgroup, a or b part.
// Rotation rot applies tau_{5^(rot+1)}: key index rot, permutation row rot+1.
//   acc_p = -sum_j D[j][p][perm(s)] * key[rot][j][ab][p][s]       (p = q1,q2,bs)
//   b-part adds bhat[p][perm(s)]  (tau(b) in the NTT domain)
//   c = INTT_bs(acc_bs) in [0,bsMod);  out_p = (acc_p - NTT_p(c)) * bsMod^-1
// out goes to ctw[k_base + rot][ab][p][z].  Rotation `mid` also stores its
// un-switched NTT-domain accumulators in mid_out[ab][p] (evalAutoSRNSFirst's
// intermediate).  twosteps.cpp:937-1204.
// The three digit planes D[0..2][p] of a prime are staged in LDS with
// coalesced loads and permuted from there.  Jobs: nrot * 2 * nq, rotation
// fastest, then a|b, then query (one workgroup each, 1-D grid).
// one rotation's key product for prime p in the rotation's slot order y
// (the order the keys are stored in): for the 16 slots
// y = 1024g + x0 + i (e = 4g + i) of this thread,
//   a[e] = -sum_j D[j][p][y] * key'[j][ab][p][y]   (+ bhat[p][y] for the b part)
// with D and bhat read in natural order; slot x of the rotation is then
// a[perm(x)]
__device__ __forceinline__ void baby_key_mult(uint32_t* a, const u32x4 (&kv)[3][4], const u32x4 (&dv)[3][4],
                                              const u32x4 (&bv)[4], int ab, const PrimeTab& P) {
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint64_t t = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) t += (uint64_t)dv[j][g][i] * kv[j][g][i];
            uint32_t r = neg_mod(mont_reduce(t, P.q, P.mont_qinv), P.q);
            if (ab) r = add_mod(r, bv[g][i], P.q);
            a[4 * g + i] = r;
        }
}

// output of a modulus switch: s[0], s[1] hold NTT_q1(c), NTT_q2(c);
// out_p = (acc_p - NTT_p(c)) * bsMod^-1 -> ctw[k][ab][p], acc_p from q12(e, p)
template <class Q12>
__device__ __forceinline__ void baby_switch_finish(Q12 q12, uint32_t (*s)[kPadN], uint32_t* ctw, int k, int ab,
                                                   int tid, const Tabs& T) {
    const int x0 = 4 * tid;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        uint32_t o1[4], o2[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = 4 * g + i, x = g * 1024 + x0 + i;
            o1[i] = mul_shoup(sub_mod(q12(e, 0), s[0][lds_pad(x)], kQ1), T.bsinv_q[0], T.bsinv_q_sh[0], kQ1);
            o2[i] = mul_shoup(sub_mod(q12(e, 1), s[1][lds_pad(x)], kQ2), T.bsinv_q[1], T.bsinv_q_sh[1], kQ2);
        }
        *reinterpret_cast<uint4*>(ctw + ((size_t)(k * 2 + ab) * 2 + 0) * kN + g * 1024 + x0) = make_uint4(o1[0], o1[1], o1[2], o1[3]);
        *reinterpret_cast<uint4*>(ctw + ((size_t)(k * 2 + ab) * 2 + 1) * kN + g * 1024 + x0) = make_uint4(o2[0], o2[1], o2[2], o2[3]);
    }
}

// modulus switch of one rotation and its store: c = INTT_bs(acc_bs),
// out_p = (acc_p - NTT_p(c)) * bsMod^-1 -> ctw[k][ab][p]; acc_bs in
// registers, acc_p (p = q1, q2) from q12(e, p); uses LDS planes 0 and 1
template <class Q12>
__device__ __forceinline__ void baby_switch_core(const uint32_t (&acc_bs)[kN / kNttThreads], Q12 q12,
                                                 uint32_t (*s)[kPadN], uint32_t* ctw, int k, int ab, int tid,
                                                 const Tabs& T) {
    constexpr int PER = kN / kNttThreads;
    const int x0 = 4 * tid;
#pragma unroll
    for (int e = 0; e < PER; ++e) s[0][lds_pad((e >> 2) * 1024 + x0 + (e & 3))] = acc_bs[e];
    __syncthreads();
    // c in [0, bsMod) is below q1 and q2: the same words are c mod q1, q2
    ntt_inv_r16(s[0], T.p[2]);
#pragma unroll
    for (int e = 0; e < PER; ++e) s[1][lds_pad(tid + e * kNttThreads)] = s[0][lds_pad(tid + e * kNttThreads)];
    __syncthreads();
    ntt_fwd_r16x2(s[0], T.p[0], s[1], T.p[1]);
    baby_switch_finish(q12, s, ctw, k, ab, tid, T);
    __syncthreads();
}

// two modulus switches whose bsMod inverse transforms run interleaved:
// rotation kA from accA (registers, all three residues), rotation kB from
// accB_bs (registers) with its q1, q2 residues from qB(e, p); uses LDS planes
// 0 and 1.  Same words as two baby_switch_core calls.
template <class QB>
__device__ __forceinline__ void baby_switch_pair(const uint32_t (&accA)[3][kN / kNttThreads],
                                                 const uint32_t (&accB_bs)[kN / kNttThreads], QB qB,
                                                 uint32_t (*s)[kPadN], uint32_t* ctw, int kA, int kB, int ab,
                                                 int tid, const Tabs& T) {
    constexpr int PER = kN / kNttThreads;
    const int x0 = 4 * tid;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        s[0][lds_pad((e >> 2) * 1024 + x0 + (e & 3))] = accA[2][e];
        s[1][lds_pad((e >> 2) * 1024 + x0 + (e & 3))] = accB_bs[e];
    }
    __syncthreads();
    ntt_inv_r16x2(s[0], T.p[2], s[1], T.p[2], tid);
    uint32_t cB[PER];   // c of rotation kB, held while rotation kA's transforms use both planes
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = lds_pad(tid + e * kNttThreads);
        cB[e] = s[1][x];
        s[1][x] = s[0][x];
    }
    __syncthreads();
    ntt_fwd_r16x2(s[0], T.p[0], s[1], T.p[1]);
    baby_switch_finish([&](int e, int p) { return accA[p][e]; }, s, ctw, kA, ab, tid, T);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = lds_pad(tid + e * kNttThreads);
        s[0][x] = cB[e];
        s[1][x] = cB[e];
    }
    __syncthreads();
    ntt_fwd_r16x2(s[0], T.p[0], s[1], T.p[1]);
    baby_switch_finish(qB, s, ctw, kB, ab, tid, T);
    __syncthreads();
}

__device__ __forceinline__ void baby_switch_store(const uint32_t (&acc)[3][kN / kNttThreads], uint32_t (*s)[kPadN],
                                                  uint32_t* ctw, int k, int ab, int tid, const Tabs& T) {
    baby_switch_core(acc[2], [&](int e, int p) { return acc[p][e]; }, s, ctw, k, ab, tid, T);
}

// Baby-step rotations, one workgroup per (rotation, a|b, query); rotation
// rot applies tau_{5^(rot+1)} with key index rot and permutation row rot+1
// (twosteps.cpp:937-1204):
//   acc_p[x] = -sum_j D[j][p][perm(x)] * key[rot][j][ab][p][x]     (p = q1,q2,bs)
//   b-part adds bhat[p][perm(x)]  (tau(b) in the NTT domain)
//   c = INTT_bs(acc_bs) in [0,bsMod);  out_p = (acc_p - NTT_p(c)) * bsMod^-1
// Rotations rot_base .. rot_base+nrot-1 of one half, output to
// ctw[k_base + rot]; rotation `mid` also stores its un-switched NTT-domain
// accumulators in mid_out[ab][p] (evalAutoSRNSFirst's intermediate).
// The keys are stored permuted (key'[perm(x)] = key[x]), so the products
// run at y = perm(x) with the digits, bhat and keys all read as contiguous
// 16-byte vectors (keys streamed past the caches that hold the query's
// digits); one permuted LDS read per prime and slot then brings them to slot
// order.  Each thread owns 4 consecutive slots per 1024-slot group
// (e = 4g + i).  Jobs: nrot * 2 * nq, rotation fastest, then a|b, then
// query (1-D grid).
__global__ void __launch_bounds__(kNttThreads) k_baby_rot(const uint32_t* __restrict__ D,
        const uint32_t* __restrict__ bhat, const KeyPtrs* __restrict__ kp, uint32_t* __restrict__ ctw,
        uint32_t* __restrict__ mid_out, int nrot, int rot_base, int nq, int k_base, int mid, size_t D_stride,
        size_t bhat_stride, QS S, Tabs T) {
    __shared__ uint32_t s[3][kPadN];
    const int L = xcd_job(blockIdx.x, nrot * 2 * nq);
    const int rot = rot_base + L % nrot, ab = (L / nrot) & 1;
    const size_t qi = L / (2 * nrot);
    D += qi * D_stride;
    bhat += qi * bhat_stride;
    ctw += qi * S.ctw;
    if (mid_out) mid_out += qi * S.mid;
    const uint16_t* pm = kp[qi].perms + (size_t)(rot + 1) * kN;
    const uint32_t* key = kp[qi].baby + (size_t)rot * 3 * 2 * 3 * kN;
    const int tid = threadIdx.x;
    constexpr int PER = kN / kNttThreads;
    const int x0 = 4 * tid;
    uint32_t acc[3][PER];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        u32x4 kv[3][4], dv[3][4], bv[4];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                kv[j][g] = __builtin_nontemporal_load(
                    reinterpret_cast<const u32x4*>(key + ((j * 2 + ab) * 3 + p) * kN + g * 1024 + x0));
                dv[j][g] = *reinterpret_cast<const u32x4*>(D + (j * 3 + p) * kN + g * 1024 + x0);
            }
#pragma unroll
        for (int g = 0; g < 4; ++g)
            bv[g] = ab ? *reinterpret_cast<const u32x4*>(bhat + p * kN + g * 1024 + x0) : u32x4{0, 0, 0, 0};
        uint32_t a[PER];
        baby_key_mult(a, kv, dv, bv, ab, T.p[p]);
#pragma unroll
        for (int e = 0; e < PER; ++e) s[p][lds_pad((e >> 2) * 1024 + x0 + (e & 3))] = a[e];
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const uint2 v = *reinterpret_cast<const uint2*>(pm + g * 1024 + x0);
        const int ps[4] = {(int)(v.x & 0xFFFFu), (int)(v.x >> 16), (int)(v.y & 0xFFFFu), (int)(v.y >> 16)};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int p = 0; p < 3; ++p) acc[p][4 * g + i] = s[p][lds_pad(ps[i])];
    }
    __syncthreads();
    if (rot == mid) {
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int e = 0; e < PER; ++e) mid_out[(ab * 3 + p) * kN + (e >> 2) * 1024 + x0 + (e & 3)] = acc[p][e];
    }
    baby_switch_store(acc, s, ctw, k_base + rot, ab, tid, T);
}

// The first and second halves of the baby steps with one read of each key:
// workgroup (rot, a|b, query) applies key rot / permutation row rot+1 both
// to the query's digits D0 (rotation rot+1 -> ctw[1 + rot], for
// rot < N1/2-2; rotation N1/2-1, which yields the intermediate, runs
// beforehand in k_baby_rot) and to the intermediate's digits D1
// (-> ctw[N1/2 + rot]).  Per prime the keys are read once (in the
// rotation's slot order) for both products.  The query's products come
// back permuted into registers; of the intermediate's, the bsMod one does
// too and the q1, q2 ones stay in LDS planes 2, 3 in slot order y until its
// modulus switch reads them permuted.  Jobs: N1/2 * 2 * nq, rotation
// fastest, then a|b, then query (1-D grid); two workgroups per CU.
__global__ void __launch_bounds__(kNttThreads) __attribute__((amdgpu_waves_per_eu(2)))
k_baby_rot_pair(const uint32_t* __restrict__ D0, const uint32_t* __restrict__ bhat0,
                const uint32_t* __restrict__ D1, const uint32_t* __restrict__ bhat1,
                const KeyPtrs* __restrict__ kp, uint32_t* __restrict__ ctw, int n1h, int nq, QS S, Tabs T) {
    __shared__ uint32_t s[4][kPadN];
    const int L = xcd_job(blockIdx.x, n1h * 2 * nq);
    const int rot = L % n1h, ab = (L / n1h) & 1;
    const size_t qi = L / (2 * n1h);
    const bool first = rot < n1h - 2;     // uniform over the workgroup
    D0 += qi * S.D;
    D1 += qi * S.D;
    bhat0 += qi * S.bhat;
    bhat1 += qi * S.mid;
    ctw += qi * S.ctw;
    const uint16_t* pm = kp[qi].perms + (size_t)(rot + 1) * kN;
    const uint32_t* key = kp[qi].baby + (size_t)rot * 3 * 2 * 3 * kN;
    const int tid = threadIdx.x;
    constexpr int PER = kN / kNttThreads;
    const int x0 = 4 * tid;
    uint2 pv[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) pv[g] = *reinterpret_cast<const uint2*>(pm + g * 1024 + x0);
    auto perm_of = [&](int e) -> int {
        const uint32_t w = (e & 2) ? pv[e >> 2].y : pv[e >> 2].x;
        return (int)((e & 1) ? (w >> 16) : (w & 0xFFFFu));
    };
    uint32_t accA[3][PER], accBbs[PER];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        u32x4 kv[3][4], dv[3][4], bv[4];
        uint32_t a[PER];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                kv[j][g] = __builtin_nontemporal_load(
                    reinterpret_cast<const u32x4*>(key + ((j * 2 + ab) * 3 + p) * kN + g * 1024 + x0));
                dv[j][g] = *reinterpret_cast<const u32x4*>(D1 + (j * 3 + p) * kN + g * 1024 + x0);
            }
#pragma unroll
        for (int g = 0; g < 4; ++g)
            bv[g] = ab ? *reinterpret_cast<const u32x4*>(bhat1 + p * kN + g * 1024 + x0) : u32x4{0, 0, 0, 0};
        baby_key_mult(a, kv, dv, bv, ab, T.p[p]);
        uint32_t* yb = p == 2 ? s[1] : s[2 + p];
#pragma unroll
        for (int e = 0; e < PER; ++e) yb[lds_pad((e >> 2) * 1024 + x0 + (e & 3))] = a[e];
        __builtin_amdgcn_sched_barrier(0);
        if (first) {
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    dv[j][g] = *reinterpret_cast<const u32x4*>(D0 + (j * 3 + p) * kN + g * 1024 + x0);
#pragma unroll
            for (int g = 0; g < 4; ++g)
                bv[g] = ab ? *reinterpret_cast<const u32x4*>(bhat0 + p * kN + g * 1024 + x0) : u32x4{0, 0, 0, 0};
            baby_key_mult(a, kv, dv, bv, ab, T.p[p]);
#pragma unroll
            for (int e = 0; e < PER; ++e) s[0][lds_pad((e >> 2) * 1024 + x0 + (e & 3))] = a[e];
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int ps = perm_of(e);
            if (first) accA[p][e] = s[0][lds_pad(ps)];
            if (p == 2) accBbs[e] = s[1][lds_pad(ps)];
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
    }
    if (first)
        baby_switch_pair(accA, accBbs, [&](int e, int p) { return s[2 + p][lds_pad(perm_of(e))]; }, s, ctw, 1 + rot,
                         n1h + rot, ab, tid, T);
    else
        baby_switch_core(accBbs, [&](int e, int p) { return s[2 + p][lds_pad(perm_of(e))]; }, s, ctw, n1h + rot, ab,
                         tid, T);
}

// second half, step 1: INTT of the intermediate's a in prime p (blockIdx.x)
__global__ void __launch_bounds__(kNttThreads) k_baby_mid_coef(const uint32_t* __restrict__ mid,
                                                               uint32_t* __restrict__ coef, QS S, Tabs T) {
    __shared__ uint32_t s[kPadN];
    const int p = blockIdx.x;
    mid += blockIdx.y * S.mid;
    coef += blockIdx.y * S.icoef;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) s[lds_pad(x)] = mid[(0 * 3 + p) * kN + x];
    __syncthreads();
    ntt_inv_r16(s, T.p[p]);
    for (int x = threadIdx.x; x < kN; x += kNttThreads) coef[p * kN + x] = s[lds_pad(x)];
}

// second half, step 2: digits of the intermediate's a (blockIdx.x = j*3 + p)
__global__ void __launch_bounds__(kNttThreads) k_baby_mid_digits(const uint32_t* __restrict__ coef,
                                                                 uint32_t* __restrict__ D, QS S, Tabs T) {
    __shared__ uint32_t s[kPadN];
    const int j = blockIdx.x / 3, p = blockIdx.x % 3;
    coef += blockIdx.y * S.icoef;
    D += blockIdx.y * S.D;
    const uint32_t qq = T.p[p].q;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        uint64_t xq = crt_q(coef[0 * kN + x], coef[1 * kN + x], T);
        unsigned __int128 v = crt80(xq, coef[2 * kN + x], T);
        s[lds_pad(x)] = digit_residue(bsgs_digit(v, j), qq);
    }
    __syncthreads();
    ntt_fwd_r16(s, T.p[p]);
    for (int x = threadIdx.x; x < kN; x += kNttThreads) D[(j * 3 + p) * kN + x] = s[lds_pad(x)];
}

// ------------------------------------------------------------- database

// crt.cpp:643-768 restated for a slot-minor layout: one lane per NTT slot z.
// A workgroup is WAVES waves on the same 64 slots; wave w takes rows w,
// w+WAVES, ... (RPW of them) so the rotation words of a slot are fetched
// once per workgroup (L1) and every database word exactly once.  Sums are
// exact u64 (N1 * (q-1)^2 < 2^64 for N1 <= 256) and reduced once per output.
// db: [N2][N1][N] packed (x mod q1) | (x mod q2) << 32 for one block
// ctw: [N1][2 ab][2 p][N] rotation residues, slot-minor (k = 0..N1-1)
// out: [N2][2 ab][2 primes][N] residues.  grid (N/64, r, nq), 64*WAVES threads.
template <int RPW, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_db_product(const uint64_t* __restrict__ db, const uint32_t* __restrict__ ctw,
                                                    uint32_t* __restrict__ out, int n1, int n2,
                                                    size_t block_words, size_t out_block_words, QS S, Tabs T) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int z = blockIdx.x * 64 + lane;
    const int blk = blockIdx.y;
    const size_t qi = blockIdx.z;
    const uint32_t* ct = ctw + qi * S.ctw + z;
    const uint64_t* dbb = db + blk * block_words + z;
    uint32_t* ob = out + qi * S.dbout + blk * out_block_words + z;
    // RPW = N2 / WAVES for N2 >= WAVES, so every row of the wave exists; for
    // N2 < WAVES (RPW = 1) the waves past the last row have nothing to do
    if (wv >= n2) return;
    uint64_t a1[RPW], a2[RPW], b1[RPW], b2[RPW];
#pragma unroll
    for (int i = 0; i < RPW; ++i) a1[i] = a2[i] = b1[i] = b2[i] = 0;
    // the row loads of one k carry no condition, so they issue together
#pragma unroll 3
    for (int k = 0; k < n1; ++k) {
        const uint32_t* c = ct + (size_t)k * 4 * kN;
        const uint32_t ca1 = c[0], ca2 = c[kN], cb1 = c[2 * kN], cb2 = c[3 * kN];
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
            const int row = wv + WAVES * i;
            const uint64_t d = __builtin_nontemporal_load(dbb + ((size_t)row * n1 + k) * kN);
            const uint32_t d1 = (uint32_t)d, d2 = (uint32_t)(d >> 32);
            a1[i] += (uint64_t)d1 * ca1;
            a2[i] += (uint64_t)d2 * ca2;
            b1[i] += (uint64_t)d1 * cb1;
            b2[i] += (uint64_t)d2 * cb2;
        }
    }
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
        const int row = wv + WAVES * i;
        uint32_t* o = ob + (size_t)row * 4 * kN;
        o[0 * kN] = (uint32_t)(a1[i] % kQ1);
        o[1 * kN] = (uint32_t)(a2[i] % kQ2);
        o[2 * kN] = (uint32_t)(b1[i] % kQ1);
        o[3 * kN] = (uint32_t)(b2[i] % kQ2);
    }
}

// ------------------------------------------------ batched database pass
// For a batch the per-slot product out[rho][col] = sum_k D[z][rho][k] *
// C[z][k][col] (rho = block*N2 + row, col = query*2 + a|b) is a small GEMM
// per slot z and residue.  Both operands are split into four 7-bit limbs
// (residues < 2^28), every limb product is an exact int8 MFMA with int32
// accumulation, and the limb pairs are summed by diagonal
//   S_w = sum_{u+v=w} A_u B_v,   w = 0..6,   0 <= S_w < 4*127^2*N1 <= 2^24 < q
// so the residue is V = sum_w S_w * 2^(7w) mod q = mont(sum_w S_w * c_w)
// with c_w = 2^(7w) * 2^32 mod q.  Requires N1 % 64 == 0, N1 <= 256 and
// r*N2 % 16 == 0.

typedef int8_t i8x16 __attribute__((ext_vector_type(16)));
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));

// database limbs, once per database: words [rho][k][z] -> dbl [z][p][u][rho][k]
// grid (N/64, R), 256 threads
__global__ void __launch_bounds__(256) k_db_limbs(const uint64_t* __restrict__ words, int8_t* __restrict__ dbl,
                                                  int n1, int R) {
    __shared__ uint64_t tile[64][65];
    const int z0 = blockIdx.x * 64, rho = blockIdx.y;
    for (int kc = 0; kc < n1; kc += 64) {
        for (int e = threadIdx.x; e < 64 * 64; e += 256) {
            const int kk = e >> 6, zz = e & 63;
            tile[kk][zz] = words[((size_t)rho * n1 + kc + kk) * kN + z0 + zz];
        }
        __syncthreads();
        // 64 slots x 8 (p,u) x 4 chunks of 16 bytes
        for (int e = threadIdx.x; e < 64 * 8 * 4; e += 256) {
            const int c16 = e & 3, pu = (e >> 2) & 7, zz = e >> 5;
            const int p = pu >> 2, u = pu & 3;
            i8x16 v;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const uint64_t w = tile[c16 * 16 + t][zz];
                const uint32_t x = p ? (uint32_t)(w >> 32) : (uint32_t)w;
                v[t] = (int8_t)((x >> (7 * u)) & 127);
            }
            int8_t* dst = dbl + ((((size_t)(z0 + zz) * 2 + p) * 4 + u) * R + rho) * n1 + kc + c16 * 16;
            *reinterpret_cast<i8x16*>(dst) = v;
        }
        __syncthreads();
    }
}

// the inverse of k_db_limbs for rows rho0 .. rho0 + gridDim.y: database words
// [row - rho0][N1][N] (x mod q1 | x mod q2 << 32) from the limbs
// [N][2 p][4 limbs][R][N1].  grid (N/64, rows)
__global__ void __launch_bounds__(256) k_db_unlimbs(const int8_t* __restrict__ dbl, uint64_t* __restrict__ words,
                                                    int n1, int R, int rho0) {
    __shared__ uint64_t tile[64][65];
    const int z0 = blockIdx.x * 64, rho = rho0 + blockIdx.y;
    const int zz = threadIdx.x >> 2, c16 = threadIdx.x & 3;   // 64 slots x 4 chunks of 16 rotations
    for (int kc = 0; kc < n1; kc += 64) {
        uint32_t x[2][16];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
#pragma unroll
            for (int t = 0; t < 16; ++t) x[p][t] = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const i8x16 v = *reinterpret_cast<const i8x16*>(
                    dbl + ((((size_t)(z0 + zz) * 2 + p) * 4 + u) * R + rho) * n1 + kc + c16 * 16);
#pragma unroll
                for (int t = 0; t < 16; ++t) x[p][t] |= (uint32_t)(uint8_t)v[t] << (7 * u);
            }
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) tile[c16 * 16 + t][zz] = x[0][t] | ((uint64_t)x[1][t] << 32);
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 64; e += 256) {
            const int kk = e >> 6, z = e & 63;
            words[((size_t)blockIdx.y * n1 + kc + kk) * kN + z0 + z] = tile[kk][z];
        }
        __syncthreads();
    }
}

// Query limbs for the int8-MFMA database pass: the rotation words ctw
// [q][k][ab][p][z] (u32 residues < 2^28) of one MFMA sweep become 7-bit limbs
// ql [z][p][4 limbs][C][N1] (column q 2 + ab), so that every (slot, prime,
// limb) is a row-major C x N1 int8 matrix.  The workgroup (32 slots, query
// q, prime p) stages the rotation words of its 32 slots with 16-byte loads
// (each 128-byte row of 32 slots by 8 lanes) in an LDS tile [slot][row]
// (row = ab N1 + k) whose rows are XOR-swizzled by 8 (slot >> 2) within each
// slot: every ds_write_b32 of the staging and every ds_read_b128 below hits
// 64 distinct banks, and 4-row groups stay contiguous.  Then lane (slot,
// 4-row group) reads its four words in one 16-byte read, extracts limb v of
// each (bits 7v .. 7v + 6) and packs the four rotations into one dword per
// limb: a wave's store of limb v covers 64 consecutive 4-row groups, i.e.
// whole 2 N1-byte runs [q][ab][k] of the output (256 bytes at N1 128).
// Columns of queries q >= nq (padding up to C) are written as zeros.
// grid (N/32, C/2, 2 p)
template <int N1>
__global__ void __launch_bounds__(256) k_ct_limbs(const uint32_t* __restrict__ ctw, int8_t* __restrict__ ql,
                                                  int nq, int C, QS S) {
    constexpr int ROWS = 2 * N1, PER = ROWS * 8 / 256, GROUPS = ROWS / 4;
    __shared__ __attribute__((aligned(16))) uint32_t tile[32 * ROWS];
    const int z0 = blockIdx.x * 32, q = blockIdx.y, p = blockIdx.z;
    const int tid = threadIdx.x;
    auto swz = [](int slot, int row) { return slot * ROWS + (row ^ ((slot >> 2) << 3)); };
    uint4 buf[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = tid + 256 * i, zq = e & 7, row = e >> 3;
        const int ab = row / N1, k = row % N1;
        buf[i] = q < nq ? *reinterpret_cast<const uint4*>(ctw + (size_t)q * S.ctw +
                                                          ((size_t)(k * 2 + ab) * 2 + p) * kN + z0 + 4 * zq)
                        : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = tid + 256 * i, zq = e & 7, row = e >> 3;
        tile[swz(4 * zq + 0, row)] = buf[i].x;
        tile[swz(4 * zq + 1, row)] = buf[i].y;
        tile[swz(4 * zq + 2, row)] = buf[i].z;
        tile[swz(4 * zq + 3, row)] = buf[i].w;
    }
    __syncthreads();
    // (ql of the chunk is below 4 GiB: 32-bit offsets)
    const uint32_t qoff = (uint32_t)q * 2 * N1;
#pragma unroll
    for (int t = 0; t < 32 * GROUPS / 256; ++t) {
        const int P = tid + 256 * t, zz = P / GROUPS, g = P % GROUPS;
        const uint4 x = *reinterpret_cast<const uint4*>(&tile[swz(zz, 4 * g)]);
        const uint32_t base = (((uint32_t)(z0 + zz) * 2 + p) * 4 * C) * N1 + qoff + 4 * g;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            // rotation 4g + i at byte i (x < 2^28: limbs of 7 bits)
            const uint32_t o = __builtin_amdgcn_ubfe(x.x, 7 * v, 7) | (__builtin_amdgcn_ubfe(x.y, 7 * v, 7) << 8) |
                               (__builtin_amdgcn_ubfe(x.z, 7 * v, 7) << 16) | (__builtin_amdgcn_ubfe(x.w, 7 * v, 7) << 24);
            *reinterpret_cast<uint32_t*>(ql + (base + (uint32_t)v * C * N1)) = o;
        }
    }
}

struct MfmaConst {
    uint32_t cw[2][7];   // 2^(7w) * 2^32 mod q_p
};

// one slot z per workgroup (4 waves), rows r0 .. r0 + R of the database's Rt
// (a chunk of whole blocks; the output's rows are chunk-local).  The query limbs of the slot for one
// residue are staged in LDS (row pitch n1 + 16 bytes: conflict-free 16-byte
// reads); each wave streams the database limb tiles of its row tiles from
// HBM once and sweeps all column tiles.  Output tile-major:
//   mf [p][rt][ct][z][16 col][16 rho] (canonical residues).
// Lane l of a 16x16x64 MFMA holds A row l%16 / B column l%16 with 16
// k-bytes at 16*(l/16); D holds rows 4*(l/16)+j of column l%16.
template <int KS>
__global__ void __launch_bounds__(256) k_db_mfma(const int8_t* __restrict__ dbl, const int8_t* __restrict__ ql,
                                                 uint32_t* __restrict__ mf, int R, int Rt, int r0, int C, MfmaConst MC,
                                                 Tabs T) {
    extern __shared__ __attribute__((aligned(16))) int8_t lds_b[];
    constexpr int n1 = KS * 64;
    constexpr int pitch = n1 + 16;
    const int z = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int RT = R / 16, CT = C / 16;
    const int lr = lane & 15, lk = (lane >> 4) * 16;
    for (int p = 0; p < 2; ++p) {
        const PrimeTab& P = T.p[p];
        const int8_t* qsrc = ql + ((size_t)z * 2 + p) * 4 * C * n1;
        for (int e = threadIdx.x; e < 4 * C * (n1 / 16); e += 256) {
            const int row = e / (n1 / 16), c16 = e % (n1 / 16);   // row = v*C + col
            *reinterpret_cast<i8x16*>(lds_b + row * pitch + c16 * 16) =
                *reinterpret_cast<const i8x16*>(qsrc + (size_t)row * n1 + c16 * 16);
        }
        __syncthreads();
        // rows r0 .. r0 + R of the Rt database rows (one chunk of blocks)
        const int8_t* asrc = dbl + ((size_t)z * 2 + p) * 4 * Rt * n1 + (size_t)r0 * n1;
        for (int rt = wv; rt < RT; rt += 4) {
            i8x16 a[4][KS];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int ks = 0; ks < KS; ++ks)
                    a[u][ks] = *reinterpret_cast<const i8x16*>(asrc + ((size_t)u * Rt + rt * 16 + lr) * n1 + ks * 64 + lk);
            for (int ct = 0; ct < CT; ++ct) {
                i32x4 acc[7];
#pragma unroll
                for (int w = 0; w < 7; ++w) acc[w] = i32x4{0, 0, 0, 0};
#pragma unroll
                for (int v = 0; v < 4; ++v)
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks) {
                        const i8x16 b = *reinterpret_cast<const i8x16*>(lds_b + (v * C + ct * 16 + lr) * pitch + ks * 64 + lk);
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            acc[u + v] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[u][ks], b, acc[u + v], 0, 0, 0);
                    }
                uint4 o;
                uint32_t* op = reinterpret_cast<uint32_t*>(&o);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    uint64_t t = 0;
#pragma unroll
                    for (int w = 0; w < 7; ++w) t += (uint64_t)(uint32_t)acc[w][j] * MC.cw[p][w];
                    op[j] = mont_reduce(t, P.q, P.mont_qinv);
                }
                uint32_t* dst = mf + ((((size_t)(p * RT + rt) * CT + ct) * kN + z) * 16 + lr) * 16 + (lane >> 4) * 4;
                *reinterpret_cast<uint4*>(dst) = o;
            }
        }
        __syncthreads();
    }
}

// The same products written straight into the first-dimension layout
// dbout [q][b][i][ab][p][z], with no tile-major intermediate.  Coalesced
// writes need runs of consecutive slots, so a workgroup owns 16 slots and
// RTW*16 rows (8 waves, each 2 slots x RTW row tiles): the database limbs of
// its (slot, row tile) pairs stay in registers for the whole sweep and are
// read from HBM exactly once.  The query limbs of a (slot, column tile) are
// read from L2 (once per row group; the row groups of one 16-slot group run
// back to back on one XCD) by LDS-DMA into a ring of NS stages per wave, one
// stage per (column tile, slot, limb v), issued NS stages ahead of use, so
// the L2 latency hides behind the MFMAs of the stages before it.  A wave
// waits for its stage g with vmcnt(KS * stages issued after g): vector loads
// complete in order, so at most that many outstanding operations means stage
// g has landed (the wave's global stores of earlier column tiles count in
// vmcnt too; whatever their order, they can only make such a wait longer).
// For each column tile the 16 x 16 products of every (slot, row tile) go
// through an LDS tile [slot][row][col] (two of them, for alternate column
// tiles) and leave as 64-byte runs of 16 slots per (row, column), after one
// LDS-only barrier per column tile: no wait for stores or prefetches in
// flight.  The strides make the MFMA result writes and the run reads touch
// each bank at most twice per wave.
// ZMAJ selects the slot-group-major first-dimension layout
//   dbout [q][b][p][z / 16][i][ab][z % 16]
// in which each workgroup's output for one (query, block) is one contiguous
// run of n2 * 2 * 64 bytes, and each store instruction writes whole 128-byte
// lines; k_giant_fused_dma<EXT, true> reads it.  Without ZMAJ the layout is
// dbout [q][b][i][ab][p][z] (64-byte runs, one per (row, column)).
// Requires R % (16 RTW) == 0.  grid: 256 slot groups x (R / (16 RTW)) row
// groups x 2 primes, 512 threads.
constexpr int kMfmaRhoStride = 20;                    // words per staged row (16 columns + pad)
constexpr int kMfmaZStride = 32 * kMfmaRhoStride + 4; // words per staged slot (up to 32 rows)

// s_waitcnt vmcnt(n), other counters not waited on (n < 16)
template <int NV>
__device__ __forceinline__ void wait_vmcnt() { __builtin_amdgcn_s_waitcnt(0x0F70 | NV); }

// KS 16-byte LDS reads at la + 1024 ks, completed (lgkmcnt(0)) within the
// same asm statement.  Written as asm so that the compiler sees no LDS read
// of the DMA ring: with global stores in flight it would otherwise guard each
// such read with vmcnt(0), which also waits for every store and prefetch.
// The ring's own vmcnt waits order these reads after their stage's DMA.
template <int KS>
__device__ __forceinline__ void ring_read(i8x16 (&b)[KS], uint32_t la) {
    i32x4 r[KS];
    if constexpr (KS == 1) {
        asm volatile("ds_read_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=&v"(r[0]) : "v"(la) : "memory");
    } else if constexpr (KS == 2) {
        asm volatile("ds_read_b128 %0, %2\n\tds_read_b128 %1, %2 offset:1024\n\ts_waitcnt lgkmcnt(0)"
                     : "=&v"(r[0]), "=&v"(r[1]) : "v"(la) : "memory");
    } else {
        static_assert(KS == 4, "ring_read: KS in {1, 2, 4}");
        asm volatile("ds_read_b128 %0, %4\n\tds_read_b128 %1, %4 offset:1024\n\t"
                     "ds_read_b128 %2, %4 offset:2048\n\tds_read_b128 %3, %4 offset:3072\n\ts_waitcnt lgkmcnt(0)"
                     : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]) : "v"(la) : "memory");
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) b[k] = __builtin_bit_cast(i8x16, r[k]);
}

// one LDS word write as asm, for the same reason (the staging tile that
// shares the LDS with the DMA ring); completed by the next lds_barrier
__device__ __forceinline__ void lds_write_u32(uint32_t la, uint32_t v) {
    asm volatile("ds_write_b32 %0, %1" : : "v"(la), "v"(v) : "memory");
}

// The run reads of a column tile as one asm statement (all reads, then
// lgkmcnt(0)): w[i][k] = word (slot 4 qz + k, row rho0 + 8 i) of the thread's
// column, la = its (slot 4 qz, row rho0) address in the staging tile.
template <int NI>
__device__ __forceinline__ void stage_read(uint32_t (&w)[NI][4], uint32_t la) {
    static_assert(kMfmaRhoStride == 20 && kMfmaZStride == 644, "stage_read: offsets assume these strides");
    if constexpr (NI == 2) {
        asm volatile("ds_read_b32 %0, %8 offset:0\n\t"
                     "ds_read_b32 %1, %8 offset:2576\n\t"
                     "ds_read_b32 %2, %8 offset:5152\n\t"
                     "ds_read_b32 %3, %8 offset:7728\n\t"
                     "ds_read_b32 %4, %8 offset:640\n\t"
                     "ds_read_b32 %5, %8 offset:3216\n\t"
                     "ds_read_b32 %6, %8 offset:5792\n\t"
                     "ds_read_b32 %7, %8 offset:8368\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(w[0][0]), "=&v"(w[0][1]), "=&v"(w[0][2]), "=&v"(w[0][3]), "=&v"(w[1][0]), "=&v"(w[1][1]), "=&v"(w[1][2]), "=&v"(w[1][3])
                     : "v"(la) : "memory");
    } else {
        static_assert(NI == 4, "stage_read: NI in {2, 4}");
        asm volatile("ds_read_b32 %0, %16 offset:0\n\t"
                     "ds_read_b32 %1, %16 offset:2576\n\t"
                     "ds_read_b32 %2, %16 offset:5152\n\t"
                     "ds_read_b32 %3, %16 offset:7728\n\t"
                     "ds_read_b32 %4, %16 offset:640\n\t"
                     "ds_read_b32 %5, %16 offset:3216\n\t"
                     "ds_read_b32 %6, %16 offset:5792\n\t"
                     "ds_read_b32 %7, %16 offset:8368\n\t"
                     "ds_read_b32 %8, %16 offset:1280\n\t"
                     "ds_read_b32 %9, %16 offset:3856\n\t"
                     "ds_read_b32 %10, %16 offset:6432\n\t"
                     "ds_read_b32 %11, %16 offset:9008\n\t"
                     "ds_read_b32 %12, %16 offset:1920\n\t"
                     "ds_read_b32 %13, %16 offset:4496\n\t"
                     "ds_read_b32 %14, %16 offset:7072\n\t"
                     "ds_read_b32 %15, %16 offset:9648\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(w[0][0]), "=&v"(w[0][1]), "=&v"(w[0][2]), "=&v"(w[0][3]), "=&v"(w[1][0]), "=&v"(w[1][1]), "=&v"(w[1][2]), "=&v"(w[1][3]), "=&v"(w[2][0]), "=&v"(w[2][1]), "=&v"(w[2][2]), "=&v"(w[2][3]), "=&v"(w[3][0]), "=&v"(w[3][1]), "=&v"(w[3][2]), "=&v"(w[3][3])
                     : "v"(la) : "memory");
    }
}

// Workgroup barrier for LDS data only: this wave's LDS reads and writes are
// done (lgkmcnt(0)) and every wave has arrived.  Unlike __syncthreads(), whose
// release fence waits for vmcnt(0), it does not wait for the wave's global
// stores and LDS-DMA loads in flight (on gfx9 both count in vmcnt), so stores
// of one column tile and prefetches of the next keep going across it.  The
// empty asm statements keep the compiler from moving memory accesses across.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <int KS, int RTW, bool ZMAJ>
__global__ void __launch_bounds__(512) k_db_mfma_direct(const int8_t* __restrict__ dbl, const int8_t* __restrict__ ql,
                                                        uint32_t* __restrict__ dbout, int R, int Rt, int r0, int C,
                                                        int n2, int nq, QS S, MfmaConst MC, Tabs T, DbSegs SG) {
    typedef __attribute__((address_space(3))) void lds_void;
    typedef __attribute__((address_space(1))) void glb_void;
    constexpr int n1 = KS * 64;
    constexpr int NS = KS == 4 ? 2 : 4;     // ring stages per wave (KS KiB each)
    __shared__ __attribute__((aligned(16))) int8_t ring[8][NS][KS * 1024];
    __shared__ uint32_t stage[2][16 * kMfmaZStride];   // column tile ct in stage[ct & 1]
    // a launch spanning databases takes this workgroup's database, rows,
    // column tiles and queries from its segment
    int ct_lo = 0, ct_hi = C / 16, q_lo = 0, q_hi = nq;
    int L;
    if (SG.n) {
        L = xcd_job(blockIdx.x, SG.wg_end[SG.n - 1]);
        const int sg = seg_of(L, SG.wg_end, SG.n);
        dbl = SG.dbl[sg];
        R = SG.R[sg];
        Rt = SG.Rt[sg];
        r0 = SG.r0[sg];
        ct_lo = SG.ct_lo[sg];
        ct_hi = SG.ct_hi[sg];
        q_lo = SG.q_lo[sg];
        q_hi = SG.q_hi[sg];
    } else {
        L = xcd_job(blockIdx.x, (kN / 16) * (R / (16 * RTW)) * 2);
    }
    const int ngroups = R / (16 * RTW);
    const int rg = L % ngroups, p = (L / ngroups) & 1, zg = L / (2 * ngroups);
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lr = lane & 15, lk = (lane >> 4) * 16;
    const PrimeTab& P = T.p[p];
    const int G = (ct_hi - ct_lo) * 8;      // stages: g = ((ct - ct_lo) * 2 + zi) * 4 + v
    // this lane's query-limb fragment of stage g: slot 2 wv + zi, limb v,
    // column ct * 16 + lr, rotations lk .. lk + 15 of each 64-rotation chunk
    auto issue = [&](int g) {
        const int ct = ct_lo + (g >> 3), zi = (g >> 2) & 1, v = g & 3;
        const int z = zg * 16 + 2 * wv + zi;
        const int8_t* src = ql + ((size_t)z * 2 + p) * 4 * C * n1 + ((size_t)v * C + ct * 16 + lr) * n1 + lk;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            __builtin_amdgcn_global_load_lds((glb_void*)(src + ks * 64), (lds_void*)&ring[wv][g % NS][ks * 1024], 16,
                                             0, 0);
    };
    // database limbs of this wave's two slots and RTW row tiles
    i8x16 a[2][RTW][4][KS];
#pragma unroll
    for (int zi = 0; zi < 2; ++zi) {
        const int z = zg * 16 + 2 * wv + zi;
        const int8_t* asrc = dbl + ((size_t)z * 2 + p) * 4 * Rt * n1 + (size_t)r0 * n1;
#pragma unroll
        for (int t = 0; t < RTW; ++t)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int ks = 0; ks < KS; ++ks)
                    a[zi][t][u][ks] = *reinterpret_cast<const i8x16*>(
                        asrc + ((size_t)u * Rt + (rg * RTW + t) * 16 + lr) * n1 + ks * 64 + lk);
    }
    __builtin_amdgcn_sched_barrier(0);
    for (int g = 0; g < NS && g < G; ++g) issue(g);
    __builtin_amdgcn_sched_barrier(0);
    for (int ct = ct_lo; ct < ct_hi; ++ct) {
#pragma unroll
        for (int zi = 0; zi < 2; ++zi) {
            const int zl = 2 * wv + zi;
            i32x4 acc[RTW][7];
#pragma unroll
            for (int t = 0; t < RTW; ++t)
#pragma unroll
                for (int w = 0; w < 7; ++w) acc[t][w] = i32x4{0, 0, 0, 0};
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int g = ((ct - ct_lo) * 2 + zi) * 4 + v;
                const int after = G - 1 - g < NS - 1 ? G - 1 - g : NS - 1;   // stages issued after g
                __builtin_amdgcn_sched_barrier(0);
                if (after >= 3) wait_vmcnt<3 * KS>();
                else if (after == 2) wait_vmcnt<2 * KS>();
                else if (after == 1) wait_vmcnt<KS>();
                else wait_vmcnt<0>();
                __builtin_amdgcn_sched_barrier(0);
                i8x16 b[KS];
                ring_read<KS>(b, (uint32_t)(uintptr_t)(lds_void*)&ring[wv][g % NS][16 * lane]);
#pragma unroll
                for (int t = 0; t < RTW; ++t)
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            acc[t][u + v] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[zi][t][u][ks], b[ks],
                                                                                  acc[t][u + v], 0, 0, 0);
                // this stage's ring slot takes stage g + NS once its reads are done
                __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
                __builtin_amdgcn_sched_barrier(0);
                if (g + NS < G) issue(g + NS);
                __builtin_amdgcn_sched_barrier(0);
            }
            // lane: column lr, rows 4 (lane / 16) + j of row tile t
            const uint32_t la = (uint32_t)(uintptr_t)(lds_void*)&stage[ct & 1][zl * kMfmaZStride +
                                                                          4 * (lane >> 4) * kMfmaRhoStride + lr];
#pragma unroll
            for (int t = 0; t < RTW; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    uint64_t sum = 0;
#pragma unroll
                    for (int w = 0; w < 7; ++w) sum += (uint64_t)(uint32_t)acc[t][w][j] * MC.cw[p][w];
                    lds_write_u32(la + 4 * (t * 16 + j) * kMfmaRhoStride, mont_reduce(sum, P.q, P.mont_qinv));
                }
        }
        // every wave's products of column tile ct are in stage[ct & 1]; that
        // buffer is rewritten at column tile ct + 2, after every wave has
        // passed the barrier of ct + 1 and so finished its run reads of ct
        lds_barrier();
        // runs: (row, column) pairs, four lanes per run of 16 slots; thread t
        // takes slots 4 (t & 3) .. + 3 of column (t >> 2) & 15 in rows
        // (t >> 6) + 8 i
        {
            constexpr int NI = RTW * 2;
            const int qz = threadIdx.x & 3, col = (threadIdx.x >> 2) & 15, rho0 = threadIdx.x >> 6;
            const int cg = ct * 16 + col, q = cg >> 1, ab = cg & 1;
            uint32_t w[NI][4];
            stage_read<NI>(w, (uint32_t)(uintptr_t)(lds_void*)&stage[ct & 1][(4 * qz) * kMfmaZStride +
                                                                              rho0 * kMfmaRhoStride + col]);
            if (q >= q_lo && q < q_hi) {
#pragma unroll
                for (int i = 0; i < NI; ++i) {
                    const int rhog = (rg * RTW) * 16 + rho0 + 8 * i, blk = rhog / n2, row = rhog % n2;
                    const size_t o = ZMAJ ? ((size_t)blk * 2 + p) * 2 * n2 * kN + (size_t)zg * 32 * n2 + row * 32 +
                                                ab * 16 + 4 * qz
                                          : ((size_t)(blk * n2 + row) * 4 + ab * 2 + p) * kN + zg * 16 + 4 * qz;
                    *reinterpret_cast<uint4*>(dbout + q * S.dbout + o) = make_uint4(w[i][0], w[i][1], w[i][2], w[i][3]);
                }
            }
        }
    }
}

// tile-major MFMA output -> dbout [q][b][i][ab][p][z] for the nq queries of
// a chunk (dbout points at the chunk's first query).  32 slots of one tile
// per workgroup: one 32 KiB contiguous read (16-byte loads, all issued before
// the first LDS write), 128-byte runs per written plane.  The LDS tile
// [cr][slot ^ ((cr >> 2) & 31)] is conflict-free for both phases.
// grid (N/32, RT*CT, 2 p)
__global__ void __launch_bounds__(256) k_mfma_scatter(const uint32_t* __restrict__ mf, uint32_t* __restrict__ dbout,
                                                      int R, int C, int n2, int nq, QS S) {
    __shared__ uint32_t tile[256 * 32];
    const int z0 = blockIdx.x * 32, t = blockIdx.y, p = blockIdx.z;
    const int RT = R / 16, CT = C / 16;
    const int rt = t / CT, ct = t % CT;
    const uint4* src = reinterpret_cast<const uint4*>(mf + (((size_t)(p * RT + rt) * CT + ct) * kN + z0) * 256);
    uint4 buf[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) buf[i] = src[threadIdx.x + 256 * i];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int e = threadIdx.x + 256 * i, zz = e >> 6, c4 = e & 63;   // words cr = 4*c4 .. 4*c4+3 of slot zz
        uint32_t* d = tile + 4 * c4 * 32 + (zz ^ (c4 & 31));
        d[0] = buf[i].x;
        d[32] = buf[i].y;
        d[64] = buf[i].z;
        d[96] = buf[i].w;
    }
    __syncthreads();
    const int zz = threadIdx.x & 31;
#pragma unroll 4
    for (int cr = threadIdx.x >> 5; cr < 256; cr += 8) {   // cr = col_in * 16 + rho_in
        const int col = ct * 16 + (cr >> 4), rho = rt * 16 + (cr & 15);
        const int q = col >> 1, ab = col & 1;
        const uint32_t val = tile[cr * 32 + (zz ^ ((cr >> 2) & 31))];
        if (q < nq) {
            const int b = rho / n2, i = rho % n2;
            dbout[q * S.dbout + ((size_t)(b * n2 + i) * 4 + ab * 2 + p) * kN + z0 + zz] = val;
        }
    }
}

// ------------------------------------------------------------ giant steps

// INTT of row c >= 1 (a part) of every block: grid (N2-1, r, 2 primes)
// dbout: [r][N2][2][2][N]  ->  gcoef: [r][N2][2][N]
__global__ void __launch_bounds__(kNttThreads) k_giant_coef(const uint32_t* __restrict__ dbout,
                                                            uint32_t* __restrict__ gcoef, int n2, QS S, Tabs T) {
    __shared__ uint32_t s[kPadN];
    const int c = blockIdx.x + 1, blk = blockIdx.y, p = blockIdx.z & 1;
    const size_t qi = blockIdx.z >> 1;
    dbout += qi * S.dbout;
    gcoef += qi * S.gcoef;
    const uint32_t* src = dbout + (((size_t)blk * n2 + c) * 4 + p) * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) s[lds_pad(x)] = src[x];
    __syncthreads();
    ntt_inv_r16(s, T.p[p]);
    uint32_t* dst = gcoef + (((size_t)blk * n2 + c) * 2 + p) * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) dst[x] = s[lds_pad(x)];
}

// Key switch of row c (twosteps.cpp:599-677), prime p: grid (N2-1, r, 2)
//   a' = -(d0*K0.a + d1*K1.a),  b' = perm_g(bhat) - (d0*K0.b + d1*K1.b)
// with d = decompose(tau_g(a)) (ell 2, PP 2^20, Bg 2^18), g = 5^(N1*c).
// gks: [r][N2][2 ab][2 p][N] (row 0 unused)
__global__ void __launch_bounds__(kNttThreads) k_giant_ks(const uint32_t* __restrict__ dbout,
        const uint32_t* __restrict__ gcoef, const KeyPtrs* __restrict__ kp, uint32_t* __restrict__ gks,
        int n1, int n2, QS S, Tabs T) {
    __shared__ uint32_t s0[kPadN];
    __shared__ uint32_t s1[kPadN];
    const int c = blockIdx.x + 1, blk = blockIdx.y, p = blockIdx.z & 1;
    const size_t qi = blockIdx.z >> 1;
    dbout += qi * S.dbout;
    gcoef += qi * S.gcoef;
    gks += qi * S.gks;
    const uint32_t* gkeys = kp[qi].giant;
    const PrimeTab& P = T.p[p];
    // g = 5^(N1*c); 5 has order N/2 mod 2N, so g^-1 = 5^(N/2 - N1*c mod N/2)
    const uint32_t e = (uint32_t)(n1 * c) % (kN / 2);
    const uint32_t g = pow5_mod2n(e);
    const uint32_t ginv = pow5_mod2n((kN / 2 - e) % (kN / 2));
    const uint32_t* a1 = gcoef + (((size_t)blk * n2 + c) * 2 + 0) * kN;
    const uint32_t* a2 = a1 + kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        uint64_t v = auto_coeff(a1, a2, x, ginv, T);
        s0[lds_pad(x)] = digit_residue(tfhe_digit<18, 20, 2>(v, 0), P.q);
        s1[lds_pad(x)] = digit_residue(tfhe_digit<18, 20, 2>(v, 1), P.q);
    }
    __syncthreads();
    ntt_fwd_r16x2(s0, P, s1, P);
    const uint32_t* key = gkeys + (size_t)(c - 1) * 2 * 2 * 2 * kN;
    const uint32_t* bh = dbout + (((size_t)blk * n2 + c) * 4 + 2 + p) * kN;
    uint32_t* o = gks + (((size_t)blk * n2 + c) * 4) * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        const uint64_t d0 = s0[lds_pad(x)], d1 = s1[lds_pad(x)];
        uint64_t ta = d0 * key[((0 * 2 + 0) * 2 + p) * kN + x] + d1 * key[((1 * 2 + 0) * 2 + p) * kN + x];
        uint64_t tb = d0 * key[((0 * 2 + 1) * 2 + p) * kN + x] + d1 * key[((1 * 2 + 1) * 2 + p) * kN + x];
        o[(0 * 2 + p) * kN + x] = neg_mod(mont_reduce(ta, P.q, P.mont_qinv), P.q);
        uint32_t rb = neg_mod(mont_reduce(tb, P.q, P.mont_qinv), P.q);
        o[(1 * 2 + p) * kN + x] = add_mod(rb, bh[perm_slot(x, g, T)], P.q);
    }
}

// --------------------------------------------------- external product

// first-dimension result of each block (row 0 + key-switched rows), NTT form,
// written to fd (optional) and INTT'd to coefficients: grid (r, 2 ab, 2 p)
__global__ void __launch_bounds__(kNttThreads) k_ext_coef(const uint32_t* __restrict__ dbout,
        const uint32_t* __restrict__ gks, uint32_t* __restrict__ fd, uint32_t* __restrict__ ecoef, int n2, QS S,
        Tabs T) {
    __shared__ uint32_t s[kPadN];
    const int blk = blockIdx.x, ab = blockIdx.y, p = blockIdx.z & 1;
    const size_t qi = blockIdx.z >> 1;
    dbout += qi * S.dbout;
    gks += qi * S.gks;
    if (fd) fd += qi * S.fd;
    ecoef += qi * S.ecoef;
    const uint32_t q = T.p[p].q;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        uint32_t v = dbout[(((size_t)blk * n2 + 0) * 4 + ab * 2 + p) * kN + x];
        // independent loads: unrolled so that they are in flight together
#pragma unroll 16
        for (int c = 1; c < n2; ++c) v = add_mod(v, gks[(((size_t)blk * n2 + c) * 4 + ab * 2 + p) * kN + x], q);
        s[lds_pad(x)] = v;
        if (fd) fd[(((size_t)blk * 2 + ab) * 2 + p) * kN + x] = v;
    }
    __syncthreads();
    ntt_inv_r16(s, T.p[p]);
    for (int x = threadIdx.x; x < kN; x += kNttThreads) ecoef[(((size_t)blk * 2 + ab) * 2 + p) * kN + x] = s[lds_pad(x)];
}

// first-dimension sum for few queries, one lane per word (grid-stride):
// fd[blk][ab][p][x] =
// dbout row 0 + sum over c >= 1 of the key-switched rows gks[blk][c]
// (NTT form; the sum k_ext_coef forms, spread over the whole GPU)
__global__ void __launch_bounds__(256) k_fd_sum(const uint32_t* __restrict__ dbout, const uint32_t* __restrict__ gks,
                                                uint32_t* __restrict__ fd, int r, int n2, int nq, QS S, Tabs T) {
    const size_t per = (size_t)r * 4 * kN, total = per * nq;
    for (size_t gidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; gidx < total;
         gidx += (size_t)gridDim.x * blockDim.x) {
        const size_t qi = gidx / per, idx = gidx % per;
        const size_t blk = idx / (4 * kN), w = idx % (4 * kN);   // w = (ab * 2 + p) * N + x
        const uint32_t q = T.p[(w / kN) & 1].q;
        const uint32_t* d = dbout + qi * S.dbout + blk * n2 * 4 * kN + w;
        const uint32_t* gk = gks + qi * S.gks + blk * n2 * 4 * kN + w;
        uint32_t v = d[0];
#pragma unroll 16
        for (int c = 1; c < n2; ++c) v = add_mod(v, gk[(size_t)c * 4 * kN], q);
        fd[qi * S.fd + idx] = v;
    }
}

// INTT of each first-dimension plane: fd -> ecoef, grid (r, 2 ab, 2 p * nq)
__global__ void __launch_bounds__(kNttThreads) k_fd_intt(const uint32_t* __restrict__ fd, uint32_t* __restrict__ ecoef,
                                                         QS S, Tabs T) {
    __shared__ uint32_t s[kPadN];
    const int blk = blockIdx.x, ab = blockIdx.y, p = blockIdx.z & 1;
    const size_t qi = blockIdx.z >> 1;
    const size_t off = (((size_t)blk * 2 + ab) * 2 + p) * kN;
    const uint32_t* src = fd + qi * S.fd + off;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) s[lds_pad(x)] = src[x];
    __syncthreads();
    ntt_inv_r16(s, T.p[p]);
    uint32_t* dst = ecoef + qi * S.ecoef + off;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) dst[x] = s[lds_pad(x)];
}

// Batched giant steps and first-dimension sum, one block of one query per
// workgroup, both primes: k_giant_coef + k_giant_ks + k_ext_coef in one pass
// that keeps every intermediate in LDS and registers.  For each c >= 1
// (twosteps.cpp:599-677) the a-row is INTT'd in LDS, tau_g (g = 5^(N1*c)) is
// read back from LDS, decomposed (ell 2, PP 2^20, Bg 2^18) and NTT'd, and the
// key-switched row is accumulated in registers; row 0 is added at the end.
// Writes fd (optional, NTT form) and ecoef (coefficients) as k_ext_coef does.
// 512 threads: the two 256-thread halves run the transforms of the two
// primes side by side, and every thread holds 8 slots of the accumulators.
// One workgroup per (query, block), r*nq of them; when that count is a
// multiple of 8 the ids are dealt so that the r blocks of a query run on one
// XCD (the hardware hands consecutive workgroups to the 8 XCDs in turn), where
// they share the query's giant-step keys in L2.  Scheduling barriers after
// each phase keep the compiler from hoisting one phase's loads into the
// previous one, which would raise register pressure past the 256-VGPR budget.
constexpr int kFusedThreads = 512;
__global__ void __launch_bounds__(kFusedThreads) k_giant_fused(const uint32_t* __restrict__ dbout,
        const KeyPtrs* __restrict__ kp, uint32_t* __restrict__ fd, uint32_t* __restrict__ ecoef, int r, int nq,
        int n1, int n2, QS S, Tabs T) {
    __shared__ uint32_t s[6][kPadN];     // 0..3 transforms, 4..5 the b-row of step c
    const int L = xcd_job(blockIdx.x, r * nq);
    const int blk = L % r;
    const size_t qi = L / r;
    dbout += qi * S.dbout;
    if (fd) fd += qi * S.fd;
    ecoef += qi * S.ecoef;
    const uint32_t* gkeys = kp[qi].giant;
    const int tid = threadIdx.x;
    const int half = __builtin_amdgcn_readfirstlane(tid >> 8), ht = tid & 255;
    constexpr int PER = kN / kFusedThreads;
    const uint32_t q1 = T.p[0].q, q2 = T.p[1].q;
    uint32_t acc[2][2][PER];    // [ab][p][e], slot x = tid + 512 e
#pragma unroll
    for (int e = 0; e < PER; ++e) acc[0][0][e] = acc[0][1][e] = acc[1][0][e] = acc[1][1][e] = 0;
    for (int c = 1; c < n2; ++c) {
        const uint32_t* row = dbout + ((size_t)blk * n2 + c) * 4 * kN;   // [ab][p][N]
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int x = tid + e * kFusedThreads;
            s[0][lds_pad(x)] = row[x];
            s[1][lds_pad(x)] = row[kN + x];
            s[4][lds_pad(x)] = row[2 * kN + x];
            s[5][lds_pad(x)] = row[3 * kN + x];
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        ntt_inv_r16(s[half], T.p[half], ht);
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t ex = (uint32_t)(n1 * c) % (kN / 2);
        const uint32_t g = pow5_mod2n(ex);
        const uint32_t ginv = pow5_mod2n((kN / 2 - ex) % (kN / 2));
        // digits of tau_g(a): mod q2 straight into planes 2, 3 (not read
        // here); mod q1 held until the gathers from planes 0, 1 are done
        uint32_t dg[2][PER];
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int x = tid + e * kFusedThreads;
            const uint64_t v = auto_coeff_lds(s[0], s[1], x, ginv, T);
            const int32_t d0 = tfhe_digit<18, 20, 2>(v, 0), d1 = tfhe_digit<18, 20, 2>(v, 1);
            dg[0][e] = digit_residue(d0, q1);
            dg[1][e] = digit_residue(d1, q1);
            s[2][lds_pad(x)] = digit_residue(d0, q2);
            s[3][lds_pad(x)] = digit_residue(d1, q2);
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int x = tid + e * kFusedThreads;
            s[0][lds_pad(x)] = dg[0][e];
            s[1][lds_pad(x)] = dg[1][e];
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        ntt_fwd_r16x2_same(s[2 * half], s[2 * half + 1], T.p[half], ht);
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t* key = gkeys + (size_t)(c - 1) * 2 * 2 * 2 * kN;   // [j][ab][p][N]
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int x = tid + e * kFusedThreads;
            const int px = perm_slot(x, g, T);
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const PrimeTab& P = T.p[p];
                const uint64_t d0 = s[2 * p][lds_pad(x)], d1 = s[2 * p + 1][lds_pad(x)];
                const uint64_t ta = d0 * key[((0 * 2 + 0) * 2 + p) * kN + x] + d1 * key[((1 * 2 + 0) * 2 + p) * kN + x];
                const uint64_t tb = d0 * key[((0 * 2 + 1) * 2 + p) * kN + x] + d1 * key[((1 * 2 + 1) * 2 + p) * kN + x];
                acc[0][p][e] = sub_mod(acc[0][p][e], mont_reduce(ta, P.q, P.mont_qinv), P.q);
                const uint32_t bb = add_mod(acc[1][p][e], s[4 + p][lds_pad(px)], P.q);
                acc[1][p][e] = sub_mod(bb, mont_reduce(tb, P.q, P.mont_qinv), P.q);
            }
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
    }
    const uint32_t* row0 = dbout + (size_t)blk * n2 * 4 * kN;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = tid + e * kFusedThreads;
#pragma unroll
        for (int ab = 0; ab < 2; ++ab)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const uint32_t v = add_mod(acc[ab][p][e], row0[(ab * 2 + p) * kN + x], T.p[p].q);
                s[ab * 2 + p][lds_pad(x)] = v;
                if (fd) fd[(((size_t)blk * 2 + ab) * 2 + p) * kN + x] = v;
            }
    }
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
    ntt_inv_r16x2(s[2 * half], T.p[0], s[2 * half + 1], T.p[1], ht);    // half = ab
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = tid + e * kFusedThreads;
#pragma unroll
        for (int t = 0; t < 4; ++t) ecoef[((size_t)blk * 4 + t) * kN + x] = s[t][lds_pad(x)];
    }
}

// Planes of a giant-step row ([ab][p][N], 4096 words each) copied from
// global memory straight into padded LDS planes dst[0..np-1] by the wave-wide
// LDS-DMA load: lane l of an instruction moves word x0 + l to LDS address
// M0 + 4 l, and 64 consecutive words starting at a multiple of 64 are
// consecutive in the padded layout.  64 instructions per plane, dealt to the
// 8 waves; the copy holds no VGPRs and completes under vmcnt.
template <int NP>
__device__ __forceinline__ void dma_planes(uint32_t (*dst)[kPadN], const uint32_t* __restrict__ src, int tid) {
    typedef __attribute__((address_space(3))) void lds_void;
    typedef __attribute__((address_space(1))) void glb_void;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
#pragma unroll
    for (int t = 0; t < NP * kN / 64 / (kFusedThreads / 64); ++t) {   // a trip count fixed at compile time
        const int b = wv + t * (kFusedThreads / 64), j = b >> 6, x0 = (b & 63) * 64;
        __builtin_amdgcn_global_load_lds((glb_void*)(src + j * kN + x0 + lane), (lds_void*)(&dst[j][lds_pad(x0)]), 4, 0,
                                         0);
    }
}

// The same for planes read only by gathers, kept unpadded in LDS: the 16-byte
// form of the copy moves 256 consecutive words per instruction (lane l to
// M0 + 16 l), 16 instructions per plane.
template <int NP>
__device__ __forceinline__ void dma_planes_flat(uint32_t (*dst)[kPadN], const uint32_t* __restrict__ src, int tid) {
    typedef __attribute__((address_space(3))) void lds_void;
    typedef __attribute__((address_space(1))) void glb_void;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
#pragma unroll
    for (int t = 0; t < NP * kN / 256 / (kFusedThreads / 64); ++t) {
        const int b = wv + t * (kFusedThreads / 64), j = b >> 4, x0 = (b & 15) * 256;
        __builtin_amdgcn_global_load_lds((glb_void*)(src + j * kN + x0 + 4 * lane), (lds_void*)(&dst[j][x0]), 16, 0,
                                         0);
    }
}

// The same copies of the planes (row i, part ab, primes 0..np-1) of one
// block from the slot-group-major layout blk [p][z / 16][i][ab][z % 16]
// (n2 rows): each lane gathers its own words, 16 consecutive slots are
// contiguous, so every 64-byte source run is read whole.
template <int NP>
__device__ __forceinline__ void dma_planes_z(uint32_t (*dst)[kPadN], const uint32_t* __restrict__ blk, int i, int ab,
                                             int n2, int tid) {
    typedef __attribute__((address_space(3))) void lds_void;
    typedef __attribute__((address_space(1))) void glb_void;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
#pragma unroll
    for (int t = 0; t < NP * kN / 64 / (kFusedThreads / 64); ++t) {
        const int b = wv + t * (kFusedThreads / 64), j = b >> 6, x0 = (b & 63) * 64, z = x0 + lane;
        const uint32_t* src = blk + (size_t)j * 2 * n2 * kN + (size_t)(z >> 4) * 32 * n2 + i * 32 + ab * 16 + (z & 15);
        __builtin_amdgcn_global_load_lds((glb_void*)src, (lds_void*)(&dst[j][lds_pad(x0)]), 4, 0, 0);
    }
}

template <int NP>
__device__ __forceinline__ void dma_planes_flat_z(uint32_t (*dst)[kPadN], const uint32_t* __restrict__ blk, int i,
                                                  int ab, int n2, int tid) {
    typedef __attribute__((address_space(3))) void lds_void;
    typedef __attribute__((address_space(1))) void glb_void;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
#pragma unroll
    for (int t = 0; t < NP * kN / 256 / (kFusedThreads / 64); ++t) {
        const int b = wv + t * (kFusedThreads / 64), j = b >> 4, x0 = (b & 15) * 256, z = x0 + 4 * lane;
        const uint32_t* src = blk + (size_t)j * 2 * n2 * kN + (size_t)(z >> 4) * 32 * n2 + i * 32 + ab * 16 + (z & 15);
        __builtin_amdgcn_global_load_lds((glb_void*)src, (lds_void*)(&dst[j][x0]), 16, 0, 0);
    }
}

// s_waitcnt vmcnt(0): this wave's global loads, LDS-DMA copies included, done
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// k_giant_fused with the rows prefetched: while row c is decomposed,
// transformed and multiplied, row c+1 (row 0 after the last) is already on
// its way into LDS planes 4..7 by LDS-DMA: its b-planes as soon as every wave
// has gathered row c's (after the first pass of the inverse transform), its
// a-planes once row c's digits are out.  The b-row of step c is added to the
// accumulator (by the slot permutation of tau_g) as soon as it is in LDS, and
// the a-row is inverse-transformed in place in planes 4, 5, so the four digit
// planes 0..3 are written in one step.  Planes: 0..3 digits [p][j] (0 = d0
// mod q1, 1 = d1 mod q1, 2 = d0 mod q2, 3 = d1 mod q2), 4..7 the row [ab][p]
// (the b-planes 6, 7 unpadded: they are only gathered).  Each thread owns the accumulator slots 4 (tid + 512 h) + i (h = 0, 1,
// i = 0..3), so keys, row 0 and fd move as 16-byte vectors.  The key switch
// of step c runs during the inverse transform of row c+1 (its digits are
// still in planes 0..3 then), the last one after the loop.  Same words as
// k_giant_fused: every update is exact mod q.
// ZMAJ: dbout in the slot-group-major layout that k_db_mfma_direct<KS, RTW, true>
// writes (rows gathered by dma_planes_z / dma_planes_flat_z).
// With EXT the external product of the block (lwe.cpp:761-808, what k_ext_digit
// and k_ext_sum compute) follows in the same workgroup, on the coefficient
// planes still in LDS: the signed digits (ell 2, PP 2^20, Bg 2^18) of a and b,
// their transforms in both primes (half h takes prime h), and the RGSW
// products summed over the four rows with one Montgomery reduction (four
// products below 2^58 < q * 2^32), written NTT-domain to `out` with the ext
// stride; without EXT `out` receives the coefficient planes (ecoef).
template <bool EXT, bool ZMAJ>
__global__ void __launch_bounds__(kFusedThreads) k_giant_fused_dma(const uint32_t* __restrict__ dbout,
        const KeyPtrs* __restrict__ kp, uint32_t* __restrict__ fd, uint32_t* __restrict__ ecoef, int r, int nq,
        int n1, int n2, QS S, Tabs T, int b0, GiantSegs GS) {
    // the planes 0..3 (digits), 4, 5 (a-row) and 6, 7 (b-row) as three LDS
    // objects: the compiler then knows that reading one while a copy into
    // another is in flight needs no wait
    __shared__ __attribute__((aligned(16))) uint32_t sd[4][kPadN];
    __shared__ __attribute__((aligned(16))) uint32_t sa[2][kPadN];
    __shared__ __attribute__((aligned(16))) uint32_t sb[2][kPadN];
    // plane i of the eight (i known at compile time where it is used)
    auto plane = [&](int i) -> uint32_t* { return i < 4 ? sd[i] : i < 6 ? sa[i - 4] : sb[i - 6]; };
    // r: blocks of this launch (dbout holds their rows); b0: index of the
    // first of them in the outputs fd / ecoef / ext.  A launch spanning
    // databases takes the query, the block count and the output buffer from
    // the workgroup's segment.
    int blk;
    size_t qi;
    if (GS.n) {
        int L = xcd_job(blockIdx.x, GS.wg_end[GS.n - 1]);
        const int sg = seg_of(L, GS.wg_end, GS.n);
        blk = L % GS.nb[sg];
        qi = GS.q_lo[sg] + L / GS.nb[sg];
        ecoef = GS.out[sg];
    } else {
        const int L = xcd_job(blockIdx.x, r * nq);
        blk = L % r;
        qi = L / r;
    }
    const int ob = b0 + blk;
    dbout += qi * S.dbout;
    if (fd) fd += qi * S.fd;
    ecoef += qi * (EXT ? S.ext : S.ecoef);
    const uint32_t* gkeys = kp[qi].giant;
    const int tid = threadIdx.x;
    const int half = __builtin_amdgcn_readfirstlane(tid >> 8), ht = tid & 255;
    constexpr int PER = kN / kFusedThreads;
    const uint32_t q1 = T.p[0].q, q2 = T.p[1].q;
    const uint32_t* rows = dbout + (size_t)blk * n2 * 4 * kN;
    uint32_t acc[2][2][PER];    // [ab][p][4 h + i], slot 4 (tid + 512 h) + i
#pragma unroll
    for (int e = 0; e < PER; ++e) acc[0][0][e] = acc[0][1][e] = acc[1][0][e] = acc[1][1][e] = 0;
    // key switch of step cs (>= 1) for slot run h, from its NTT'd digits in planes 0..3
    auto mac = [&](int cs, int h) {
        const uint32_t* key = gkeys + (size_t)(cs - 1) * 2 * 2 * 2 * kN;   // [j][ab][p][N]
        const int x0 = 4 * (tid + kFusedThreads * h);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const PrimeTab& P = T.p[p];
            const uint4 ka0 = *reinterpret_cast<const uint4*>(key + ((0 * 2 + 0) * 2 + p) * kN + x0);
            const uint4 ka1 = *reinterpret_cast<const uint4*>(key + ((1 * 2 + 0) * 2 + p) * kN + x0);
            const uint4 kb0 = *reinterpret_cast<const uint4*>(key + ((0 * 2 + 1) * 2 + p) * kN + x0);
            const uint4 kb1 = *reinterpret_cast<const uint4*>(key + ((1 * 2 + 1) * 2 + p) * kN + x0);
            const uint32_t ka0v[4] = {ka0.x, ka0.y, ka0.z, ka0.w}, ka1v[4] = {ka1.x, ka1.y, ka1.z, ka1.w};
            const uint32_t kb0v[4] = {kb0.x, kb0.y, kb0.z, kb0.w}, kb1v[4] = {kb1.x, kb1.y, kb1.z, kb1.w};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint64_t d0 = sd[2 * p][lds_pad(x0) + i], d1 = sd[2 * p + 1][lds_pad(x0) + i];
                const uint64_t ta = d0 * ka0v[i] + d1 * ka1v[i];
                const uint64_t tb = d0 * kb0v[i] + d1 * kb1v[i];
                acc[0][p][4 * h + i] = sub_mod(acc[0][p][4 * h + i], mont_reduce(ta, P.q, P.mont_qinv), P.q);
                acc[1][p][4 * h + i] = sub_mod(acc[1][p][4 * h + i], mont_reduce(tb, P.q, P.mont_qinv), P.q);
            }
        }
    };
    // row i's a-planes into planes 4, 5 (padded), its b-planes into 6, 7 (flat)
    auto dma_a = [&](int i) {
        if (ZMAJ) dma_planes_z<2>(sa, rows, i, 0, n2, tid);
        else dma_planes<2>(sa, rows + (size_t)i * 4 * kN, tid);
    };
    auto dma_b = [&](int i) {
        if (ZMAJ) dma_planes_flat_z<2>(sb, rows, i, 1, n2, tid);
        else dma_planes_flat<2>(sb, rows + (size_t)i * 4 * kN + 2 * kN, tid);
    };
    dma_a(n2 > 1 ? 1 : 0);
    dma_b(n2 > 1 ? 1 : 0);
    for (int c = 1; c < n2; ++c) {
        const int next = c + 1 < n2 ? c + 1 : 0;
        wait_vmem();
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t ex = (uint32_t)(n1 * c) % (kN / 2);
        const uint32_t g = pow5_mod2n(ex);
        const uint32_t ginv = pow5_mod2n((kN / 2 - ex) % (kN / 2));
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int px = perm_slot(4 * (tid + kFusedThreads * (e >> 2)) + (e & 3), g, T);
            acc[1][0][e] = add_mod(acc[1][0][e], sb[0][px], q1);
            acc[1][1][e] = add_mod(acc[1][1][e], sb[1][px], q2);
        }
        __builtin_amdgcn_sched_barrier(0);
        // the key switch of step c-1 (digits still in planes 0..3) rides along
        // the inverse transform of row c (planes 4, 5): its key loads wait out
        // while the transform computes.  Both halves end before the transform's
        // second barrier, so every thread is done with planes 0..3 before the
        // digits of row c overwrite them.
        ntt_pass_lds<2, true>(sa[half], T.p[half], ht);
        if (c >= 2) mac(c - 1, 0);
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        ntt_pass_lds<1, true>(sa[half], T.p[half], ht);
        if (c >= 2) mac(c - 1, 1);
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        // every wave is past its b-gather: row c+1's b-planes are copied now.
        // With a copy in flight the compiler makes the first use of any
        // later vector load wait for the copy as well (vmcnt(0)); the next
        // ones are the forward transform's pass-1 twiddles, two thirds of a
        // step away (pass-0 twiddles are scalar loads)
        dma_b(next);
        __builtin_amdgcn_sched_barrier(0);
        ntt_pass_lds<0, true>(sa[half], T.p[half], ht);
        // from here to the top of the loop a copy is in flight: barriers for
        // LDS data only (a __syncthreads() would wait for the copy)
        lds_barrier();
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int x = tid + e * kFusedThreads;
            const uint64_t v = auto_coeff_lds(sa[0], sa[1], x, ginv, T);
            const int32_t d0 = tfhe_digit<18, 20, 2>(v, 0), d1 = tfhe_digit<18, 20, 2>(v, 1);
            sd[0][lds_pad(x)] = digit_residue(d0, q1);
            sd[1][lds_pad(x)] = digit_residue(d1, q1);
            sd[2][lds_pad(x)] = digit_residue(d0, q2);
            sd[3][lds_pad(x)] = digit_residue(d1, q2);
        }
        lds_barrier();
        __builtin_amdgcn_sched_barrier(0);
        // forward transforms of the digits; row c+1's a-planes are copied
        // once the last pass's twiddles have arrived (as above)
        ntt_pass_lds2<0, false, true>(sd[2 * half], T.p[half], sd[2 * half + 1], T.p[half], ht);
        lds_barrier();
        ntt_pass_lds2<1, false, true>(sd[2 * half], T.p[half], sd[2 * half + 1], T.p[half], ht);
        lds_barrier();
        {
            uint32_t w[15], ws[15];
            load_pass_twiddles(T.p[half].fw2, 2, ht, w, ws);
            wait_vmem();
            __builtin_amdgcn_sched_barrier(0);
            dma_a(next);
            __builtin_amdgcn_sched_barrier(0);
            ntt_pass_lds2_tw<2, false>(sd[2 * half], sd[2 * half + 1], T.p[half], ht, w, ws);
        }
        lds_barrier();
        __builtin_amdgcn_sched_barrier(0);
    }
    if (n2 > 1) {   // the last step's key switch
        mac(n2 - 1, 0);
        mac(n2 - 1, 1);
    }
    // row 0 is in planes 4..7; each thread writes planes 0..3 only at its own
    // slots, which no other thread reads until the barrier below
    wait_vmem();
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int x0 = 4 * (tid + kFusedThreads * h);
#pragma unroll
        for (int ab = 0; ab < 2; ++ab)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                uint32_t v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t r0 = ab ? sb[p][x0 + i] : sa[p][lds_pad(x0) + i];   // b-planes unpadded
                    v[i] = add_mod(acc[ab][p][4 * h + i], r0, T.p[p].q);
                    sd[ab * 2 + p][lds_pad(x0) + i] = v[i];
                }
                if (fd)
                    *reinterpret_cast<uint4*>(fd + (((size_t)ob * 2 + ab) * 2 + p) * kN + x0) =
                        make_uint4(v[0], v[1], v[2], v[3]);
            }
    }
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
    ntt_inv_r16x2(sd[2 * half], T.p[0], sd[2 * half + 1], T.p[1], ht);    // half = ab
    if (!EXT) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int x0 = 4 * (tid + kFusedThreads * h);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const uint32_t* sp = &sd[t][lds_pad(x0)];
                *reinterpret_cast<uint4*>(ecoef + ((size_t)ob * 4 + t) * kN + x0) =
                    make_uint4(sp[0], sp[1], sp[2], sp[3]);
            }
        }
        return;
    }
    // external product: digit k = 2 src + j of part src (a, b)
    __builtin_amdgcn_sched_barrier(0);
    int32_t dg[2][4][4];   // [h][k][i]
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int x0 = 4 * (tid + kFusedThreads * h);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int src = 0; src < 2; ++src) {
                const uint64_t v = crt_q(sd[2 * src][lds_pad(x0) + i], sd[2 * src + 1][lds_pad(x0) + i], T);
                dg[h][2 * src][i] = signed_digit<18, 20>(v, 0);
                dg[h][2 * src + 1][i] = signed_digit<18, 20>(v, 1);
            }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int x0 = 4 * (tid + kFusedThreads * h);
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int i = 0; i < 4; ++i) plane(2 * k + p)[lds_pad(x0) + i] = digit_residue(dg[h][k][i], T.p[p].q);
    }
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
    ntt_fwd_r16x2_same(sd[half], sd[2 + half], T.p[half], ht);    // digits 0, 1 in prime `half`
    ntt_fwd_r16x2_same(sa[half], sb[half], T.p[half], ht);    // digits 2, 3
    __builtin_amdgcn_sched_barrier(0);
    const uint32_t* rg = kp[qi].rgsw;   // [k][ab][p][N], Montgomery form
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int x0 = 4 * (tid + kFusedThreads * h);
#pragma unroll
        for (int ab = 0; ab < 2; ++ab)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const PrimeTab& P = T.p[p];
                uint64_t acc[4] = {0, 0, 0, 0};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint4 kv = *reinterpret_cast<const uint4*>(rg + ((k * 2 + ab) * 2 + p) * kN + x0);
                    const uint32_t* sp = &plane(2 * k + p)[lds_pad(x0)];
                    acc[0] += (uint64_t)sp[0] * kv.x;
                    acc[1] += (uint64_t)sp[1] * kv.y;
                    acc[2] += (uint64_t)sp[2] * kv.z;
                    acc[3] += (uint64_t)sp[3] * kv.w;
                }
                *reinterpret_cast<uint4*>(ecoef + (((size_t)ob * 2 + ab) * 2 + p) * kN + x0) =
                    make_uint4(mont_reduce(acc[0], P.q, P.mont_qinv), mont_reduce(acc[1], P.q, P.mont_qinv),
                               mont_reduce(acc[2], P.q, P.mont_qinv), mont_reduce(acc[3], P.q, P.mont_qinv));
            }
    }
}

// k_giant_fused in four LDS planes and at most 128 VGPRs, so that two
// workgroups share a CU: the b-row of a giant step is gathered from global
// memory (L2) instead of staged in LDS, and each half runs its transforms one
// at a time.  Same words as k_giant_fused.
__global__ void __launch_bounds__(kFusedThreads) __attribute__((amdgpu_waves_per_eu(4))) k_giant_fused2(const uint32_t* __restrict__ dbout,
        const KeyPtrs* __restrict__ kp, uint32_t* __restrict__ fd, uint32_t* __restrict__ ecoef, int r, int nq,
        int n1, int n2, QS S, Tabs T) {
    __shared__ uint32_t s[4][kPadN];     // the transforms; the b-row is gathered from global memory
    const int L = xcd_job(blockIdx.x, r * nq);
    const int blk = L % r;
    const size_t qi = L / r;
    dbout += qi * S.dbout;
    if (fd) fd += qi * S.fd;
    ecoef += qi * S.ecoef;
    const uint32_t* gkeys = kp[qi].giant;
    const int tid = threadIdx.x;
    const int half = __builtin_amdgcn_readfirstlane(tid >> 8), ht = tid & 255;
    constexpr int PER = kN / kFusedThreads;
    const uint32_t q1 = T.p[0].q, q2 = T.p[1].q;
    uint32_t acc[2][2][PER];    // [ab][p][e], slot x = tid + 512 e
#pragma unroll
    for (int e = 0; e < PER; ++e) acc[0][0][e] = acc[0][1][e] = acc[1][0][e] = acc[1][1][e] = 0;
    for (int c = 1; c < n2; ++c) {
        const uint32_t* row = dbout + ((size_t)blk * n2 + c) * 4 * kN;   // [ab][p][N]
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int x = tid + e * kFusedThreads;
            s[0][lds_pad(x)] = row[x];
            s[1][lds_pad(x)] = row[kN + x];
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        ntt_inv_r16(s[half], T.p[half], ht);
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t ex = (uint32_t)(n1 * c) % (kN / 2);
        const uint32_t g = pow5_mod2n(ex);
        const uint32_t ginv = pow5_mod2n((kN / 2 - ex) % (kN / 2));
        // digits of tau_g(a): mod q2 straight into planes 2, 3 (not read
        // here); mod q1 held until the gathers from planes 0, 1 are done
        uint32_t dg[2][PER];
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int x = tid + e * kFusedThreads;
            const uint64_t v = auto_coeff_lds(s[0], s[1], x, ginv, T);
            const int32_t d0 = tfhe_digit<18, 20, 2>(v, 0), d1 = tfhe_digit<18, 20, 2>(v, 1);
            dg[0][e] = digit_residue(d0, q1);
            dg[1][e] = digit_residue(d1, q1);
            s[2][lds_pad(x)] = digit_residue(d0, q2);
            s[3][lds_pad(x)] = digit_residue(d1, q2);
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int x = tid + e * kFusedThreads;
            s[0][lds_pad(x)] = dg[0][e];
            s[1][lds_pad(x)] = dg[1][e];
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        ntt_fwd_r16(s[2 * half], T.p[half], ht);
        ntt_fwd_r16(s[2 * half + 1], T.p[half], ht);
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t* key = gkeys + (size_t)(c - 1) * 2 * 2 * 2 * kN;   // [j][ab][p][N]
#pragma unroll 2
        for (int e = 0; e < PER; ++e) {
            const int x = tid + e * kFusedThreads;
            const int px = perm_slot(x, g, T);
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const PrimeTab& P = T.p[p];
                const uint64_t d0 = s[2 * p][lds_pad(x)], d1 = s[2 * p + 1][lds_pad(x)];
                const uint64_t ta = d0 * key[((0 * 2 + 0) * 2 + p) * kN + x] + d1 * key[((1 * 2 + 0) * 2 + p) * kN + x];
                const uint64_t tb = d0 * key[((0 * 2 + 1) * 2 + p) * kN + x] + d1 * key[((1 * 2 + 1) * 2 + p) * kN + x];
                acc[0][p][e] = sub_mod(acc[0][p][e], mont_reduce(ta, P.q, P.mont_qinv), P.q);
                const uint32_t bb = add_mod(acc[1][p][e], row[(2 + p) * kN + px], P.q);
                acc[1][p][e] = sub_mod(bb, mont_reduce(tb, P.q, P.mont_qinv), P.q);
            }
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
    }
    const uint32_t* row0 = dbout + (size_t)blk * n2 * 4 * kN;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = tid + e * kFusedThreads;
#pragma unroll
        for (int ab = 0; ab < 2; ++ab)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const uint32_t v = add_mod(acc[ab][p][e], row0[(ab * 2 + p) * kN + x], T.p[p].q);
                s[ab * 2 + p][lds_pad(x)] = v;
                if (fd) fd[(((size_t)blk * 2 + ab) * 2 + p) * kN + x] = v;
            }
    }
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
    ntt_inv_r16(s[2 * half], T.p[0], ht);    // half = ab
    ntt_inv_r16(s[2 * half + 1], T.p[1], ht);
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = tid + e * kFusedThreads;
#pragma unroll
        for (int t = 0; t < 4; ++t) ecoef[((size_t)blk * 4 + t) * kN + x] = s[t][lds_pad(x)];
    }
}

// lwe.cpp:761-808: digit k (0,1 from a; 2,3 from b) of decompose_rlwe
// (ell 2, PP 2^20, Bg 2^18), NTT, times RGSW row k: grid (r, 4, 2 p)
// epart: [r][4][2 ab][2 p][N]
__global__ void __launch_bounds__(kNttThreads) k_ext_digit(const uint32_t* __restrict__ ecoef,
        const KeyPtrs* __restrict__ kp, uint32_t* __restrict__ epart, QS S, Tabs T) {
    __shared__ uint32_t s[kPadN];
    const int blk = blockIdx.x, k = blockIdx.y, p = blockIdx.z & 1;
    const size_t qi = blockIdx.z >> 1;
    ecoef += qi * S.ecoef;
    epart += qi * S.epart;
    const uint32_t* rgsw = kp[qi].rgsw;
    const PrimeTab& P = T.p[p];
    const int src = k >> 1, j = k & 1;
    const uint32_t* c1 = ecoef + (((size_t)blk * 2 + src) * 2 + 0) * kN;
    const uint32_t* c2 = c1 + kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        uint64_t v = crt_q(c1[x], c2[x], T);
        s[lds_pad(x)] = digit_residue(signed_digit<18, 20>(v, j), P.q);
    }
    __syncthreads();
    ntt_fwd_r16(s, P);
    uint32_t* o = epart + (((size_t)blk * 4 + k) * 4) * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        const uint64_t d = s[lds_pad(x)];
        o[(0 * 2 + p) * kN + x] = mont_reduce(d * rgsw[((k * 2 + 0) * 2 + p) * kN + x], P.q, P.mont_qinv);
        o[(1 * 2 + p) * kN + x] = mont_reduce(d * rgsw[((k * 2 + 1) * 2 + p) * kN + x], P.q, P.mont_qinv);
    }
}

// ext[blk][ab][p] = sum_k epart[blk][k][ab][p]  (elementwise, 4*r*N lanes)
__global__ void __launch_bounds__(256) k_ext_sum(const uint32_t* __restrict__ epart, uint32_t* __restrict__ ext,
                                                 int r, int nq, QS S, Tabs T) {
    const size_t per = (size_t)r * 4 * kN;
    const size_t total = per * nq;
    for (size_t gidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; gidx < total;
         gidx += (size_t)gridDim.x * blockDim.x) {
        const size_t qi = gidx / per, idx = gidx % per;
        const uint32_t* epart_q = epart + qi * S.epart;
        const size_t blk = idx / (4 * kN);
        const size_t w = idx % (4 * kN);     // (ab, p, x)
        const int p = (int)((w / kN) & 1);
        const uint32_t q = T.p[p].q;
        uint32_t v = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) v = add_mod(v, epart_q[(blk * 4 + k) * 4 * kN + w], q);
        ext[qi * S.ext + idx] = v;
    }
}

// ---------------------------------------------------------------- packing
// one level of answer.cpp:800-835 with i pairs (j, i+j), shift X^i,
// automorphism index g = N/i + 1, key level li.  ext: [r][2 ab][2 p][N].

// (ext[j].a - ext[i+j].a * X^i) -> INTT -> pcoef[j][p]: grid (i, 2 p)
__global__ void __launch_bounds__(kNttThreads) k_pack_coef(const uint32_t* __restrict__ ext,
        uint32_t* __restrict__ pcoef, int i, QS S, Tabs T) {
    __shared__ uint32_t s[kPadN];
    const int j = blockIdx.x, p = blockIdx.y & 1;
    const size_t qi = blockIdx.y >> 1;
    ext += qi * S.ext;
    pcoef += qi * S.pcoef;
    const PrimeTab& P = T.p[p];
    const uint32_t* ej = ext + ((size_t)j * 4 + p) * kN;
    const uint32_t* ei = ext + ((size_t)(i + j) * 4 + p) * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads)
        s[lds_pad(x)] = sub_mod(ej[x], mont_mul(ei[x], xpow_m(x, i, P, T), P), P.q);
    __syncthreads();
    ntt_inv_r16(s, P);
    for (int x = threadIdx.x; x < kN; x += kNttThreads) pcoef[((size_t)j * 2 + p) * kN + x] = s[lds_pad(x)];
}

// answer.cpp:576-648 evalAuto: digit k of decompose(tau_g(a)) (ell 4, PP 0,
// Bg 2^14), NTT, times packing key: grid (i, 4, 2 p);  ppart: [i][4][2][2][N]
__global__ void __launch_bounds__(kNttThreads) k_pack_digit(const uint32_t* __restrict__ pcoef,
        const KeyPtrs* __restrict__ kp, int li, uint32_t* __restrict__ ppart, uint32_t ginv, QS S, Tabs T) {
    __shared__ uint32_t s[kPadN];
    const int j = blockIdx.x, k = blockIdx.y, p = blockIdx.z & 1;
    const size_t qi = blockIdx.z >> 1;
    pcoef += qi * S.pcoef;
    ppart += qi * S.ppart;
    const uint32_t* pkey = kp[qi].packing + (size_t)li * 4 * 2 * 2 * kN;
    const PrimeTab& P = T.p[p];
    const uint32_t* a1 = pcoef + ((size_t)j * 2 + 0) * kN;
    const uint32_t* a2 = a1 + kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        uint64_t v = auto_coeff(a1, a2, x, ginv, T);
        s[lds_pad(x)] = digit_residue(tfhe_digit<14, 0, 4>(v, k), P.q);
    }
    __syncthreads();
    ntt_fwd_r16(s, P);
    uint32_t* o = ppart + (((size_t)j * 4 + k) * 4) * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        const uint64_t d = s[lds_pad(x)];
        o[(0 * 2 + p) * kN + x] = mont_reduce(d * pkey[((k * 2 + 0) * 2 + p) * kN + x], P.q, P.mont_qinv);
        o[(1 * 2 + p) * kN + x] = mont_reduce(d * pkey[((k * 2 + 1) * 2 + p) * kN + x], P.q, P.mont_qinv);
    }
}

// One whole packing level (answer.cpp:800-835) per (pair j, query): the
// work of k_pack_coef, k_pack_digit and k_pack_combine in one workgroup of
// 512 threads, the two 256-thread halves running the transforms of the two
// primes.  a_j - a_{i+j} X^i is INTT'd in LDS, tau_g of it is read back from
// LDS and decomposed (ell 4, PP 0, Bg 2^14, answer.cpp:576-648) into eight
// LDS planes (four digits, two primes), the digits are NTT'd two at a time
// and their key products summed in registers, the first pair's products
// during the second pair's transforms; then convertTwoRLWEs
// (answer.cpp:720-739) writes next[j].  Same words as the
// three kernels (sums of Montgomery products are reduced once, an exact
// identity mod q).  Scheduling barriers keep each phase's loads in that
// phase.  1-D grid of i * nq workgroups.
__global__ void __launch_bounds__(512) k_pack_level(const uint32_t* __restrict__ ext,
        const KeyPtrs* __restrict__ kp, int li, uint32_t* __restrict__ next, int i, uint32_t g, uint32_t ginv, QS S,
        Tabs T, PackSegs PS) {
    __shared__ uint32_t s[8][kPadN];
    constexpr int PER = kN / 512;
    int L = blockIdx.x;
    const int sg = PS.n ? seg_of(L, PS.wg_end, PS.n) : 0;
    const int j = L % i;
    const size_t qi = (PS.n ? PS.q_lo[sg] : 0) + L / i;
    ext += qi * S.ext;
    next += qi * S.ext;
    const uint32_t* pkey = kp[qi].packing + (size_t)li * 4 * 2 * 2 * kN;   // [k][ab][p][N]
    const int tid = threadIdx.x;
    const int half = __builtin_amdgcn_readfirstlane(tid >> 8), ht = tid & 255;
    const uint32_t* ej = ext + (size_t)j * 4 * kN;         // [ab][p][N]
    const uint32_t* ei = ext + (size_t)(i + j) * 4 * kN;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = tid + e * 512;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const PrimeTab& P = T.p[p];
            s[p][lds_pad(x)] = sub_mod(ej[p * kN + x], mont_mul(ei[p * kN + x], xpow_m(x, i, P, T), P), P.q);
        }
    }
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
    ntt_inv_r16(s[half], T.p[half], ht);
    __builtin_amdgcn_sched_barrier(0);
    // digits of tau_g: digit k of prime p at plane D(k, p): digits 0, 1 in
    // planes 2, 3 (q1) and 4, 5 (q2), digits 2, 3 in planes 6, 7 (q1) and 0, 1
    // (q2); the latter two only once every gather from planes 0, 1 is done
    uint32_t dq2[2][PER];   // digits 2, 3 mod q2, held for planes 0, 1
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = tid + e * 512;
        const uint64_t v = auto_coeff_lds(s[0], s[1], x, ginv, T);
        int32_t d[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) d[k] = tfhe_digit<14, 0, 4>(v, k);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            s[2 + 2 * p][lds_pad(x)] = digit_residue(d[0], T.p[p].q);
            s[3 + 2 * p][lds_pad(x)] = digit_residue(d[1], T.p[p].q);
        }
        s[6][lds_pad(x)] = digit_residue(d[2], T.p[0].q);
        s[7][lds_pad(x)] = digit_residue(d[3], T.p[0].q);
        dq2[0][e] = digit_residue(d[2], T.p[1].q);
        dq2[1][e] = digit_residue(d[3], T.p[1].q);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = tid + e * 512;
        s[0][lds_pad(x)] = dq2[0][e];
        s[1][lds_pad(x)] = dq2[1][e];
    }
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
    uint32_t acc[2][2][PER];   // [ab][p][e]: sum_k digit_k * key_k
    // key products of digits k0, k0 + 1 (NTT'd, at planes pl[p][0..1]) for the slots e0 .. e0 + 3
    auto mac = [&](int k0, const int (&pl)[2][2], int e0) {
#pragma unroll
        for (int e = e0; e < e0 + PER / 2; ++e) {
            const int x = tid + e * 512;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const PrimeTab& P = T.p[p];
                const uint64_t d0 = s[pl[p][0]][lds_pad(x)], d1 = s[pl[p][1]][lds_pad(x)];
#pragma unroll
                for (int ab = 0; ab < 2; ++ab) {
                    const uint64_t t = d0 * pkey[((k0 * 2 + ab) * 2 + p) * kN + x] +
                                       d1 * pkey[(((k0 + 1) * 2 + ab) * 2 + p) * kN + x];
                    const uint32_t m = mont_reduce(t, P.q, P.mont_qinv);
                    acc[ab][p][e] = k0 ? add_mod(acc[ab][p][e], m, P.q) : m;
                }
            }
        }
    };
    constexpr int pl0[2][2] = {{2, 3}, {4, 5}}, pl1[2][2] = {{6, 7}, {0, 1}};
    // digits 0, 1: half h transforms prime h's two planes
    ntt_fwd_r16x2_same(s[2 + 2 * half], s[3 + 2 * half], T.p[half], ht);
    __builtin_amdgcn_sched_barrier(0);
    // digits 2, 3 (half 0: planes 6, 7 mod q1; half 1: planes 0, 1 mod q2), with
    // the key products of digits 0, 1 riding along the first two passes
    uint32_t* r0 = half ? s[0] : s[6];
    uint32_t* r1 = half ? s[1] : s[7];
    ntt_pass_lds2<0, false, true>(r0, T.p[half], r1, T.p[half], ht);
    mac(0, pl0, 0);
    __syncthreads();
    ntt_pass_lds2<1, false, true>(r0, T.p[half], r1, T.p[half], ht);
    mac(0, pl0, PER / 2);
    __syncthreads();
    ntt_pass_lds2<2, false, true>(r0, T.p[half], r1, T.p[half], ht);
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
    mac(2, pl1, 0);
    mac(2, pl1, PER / 2);
    __builtin_amdgcn_sched_barrier(0);
    // convertTwoRLWEs: c' = ext[i+j] * X^i + ext[j]; new = -sum + c' (+ tau_g of
    // ext[j].b - ext[i+j].b X^i for b)
    uint32_t* o = next + (size_t)j * 4 * kN;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int x = tid + e * 512;
        const int ps = perm_slot(x, g, T);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const PrimeTab& P = T.p[p];
#pragma unroll
            for (int ab = 0; ab < 2; ++ab) {
                const uint32_t* ejp = ej + (ab * 2 + p) * kN;
                const uint32_t* eip = ei + (ab * 2 + p) * kN;
                const uint32_t c = add_mod(mont_mul(eip[x], xpow_m(x, i, P, T), P), ejp[x], P.q);
                uint32_t v = neg_mod(acc[ab][p][e], P.q);
                if (ab) v = add_mod(v, sub_mod(ejp[ps], mont_mul(eip[ps], xpow_m(ps, i, P, T), P), P.q), P.q);
                o[(ab * 2 + p) * kN + x] = add_mod(v, c, P.q);
            }
        }
    }
}

// answer.cpp:720-739 convertTwoRLWEs combine: grid (i, 2 ab, 2 p)
//   c' = ext[i+j]*X^i + ext[j];  new = -sum_k ppart  (+ perm_g(ext[j].b - ext[i+j].b*X^i) for b)
//   next[j] = new + c'
__global__ void __launch_bounds__(kNttThreads) k_pack_combine(const uint32_t* __restrict__ ext,
        const uint32_t* __restrict__ ppart, uint32_t* __restrict__ next, int i, uint32_t g, QS S, Tabs T) {
    const int j = blockIdx.x, ab = blockIdx.y, p = blockIdx.z & 1;
    const size_t qi = blockIdx.z >> 1;
    ext += qi * S.ext;
    ppart += qi * S.ppart;
    next += qi * S.ext;
    const PrimeTab& P = T.p[p];
    const uint32_t* ej = ext + ((size_t)j * 4 + ab * 2 + p) * kN;
    const uint32_t* ei = ext + ((size_t)(i + j) * 4 + ab * 2 + p) * kN;
    const uint32_t* pp = ppart + ((size_t)j * 4 * 4 + ab * 2 + p) * kN;
    uint32_t* o = next + ((size_t)j * 4 + ab * 2 + p) * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        uint32_t c = add_mod(mont_mul(ei[x], xpow_m(x, i, P, T), P), ej[x], P.q);
        uint32_t t = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) t = add_mod(t, pp[(size_t)k * 4 * kN + x], P.q);
        uint32_t v = neg_mod(t, P.q);
        if (ab) {
            const int ps = perm_slot(x, g, T);
            v = add_mod(v, sub_mod(ej[ps], mont_mul(ei[ps], xpow_m(ps, i, P, T), P), P.q), P.q);
        }
        o[x] = add_mod(v, c, P.q);
    }
}

// INTT of both primes and CRT to u64 words in [0, Q): grid (2 ab, cnt)
// ext + c*ext_stride: [2 ab][2 p][N] NTT domain -> resp: [cnt][2][N] coefficient form
__global__ void __launch_bounds__(kNttThreads) k_response(const uint32_t* __restrict__ ext,
                                                          uint64_t* __restrict__ resp, size_t ext_stride, Tabs T) {
    __shared__ uint32_t s1[kPadN];
    __shared__ uint32_t s2[kPadN];
    const int ab = blockIdx.x;
    const uint32_t* e = ext + (size_t)blockIdx.y * ext_stride;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) {
        s1[lds_pad(x)] = e[(ab * 2 + 0) * kN + x];
        s2[lds_pad(x)] = e[(ab * 2 + 1) * kN + x];
    }
    __syncthreads();
    ntt_inv_r16x2(s1, T.p[0], s2, T.p[1]);
    uint64_t* o = resp + (size_t)blockIdx.y * 2 * kN + ab * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) o[x] = crt_q(s1[lds_pad(x)], s2[lds_pad(x)], T);
}

// u64 words mod Q [cnt][2][N] -> RNS planes [cnt][2 ab][2 p][N]; forward NTT
// per plane when ntt != 0: grid (cnt, 2 ab, 2 p)
__global__ void __launch_bounds__(kNttThreads) k_u64_to_rns(const uint64_t* __restrict__ in,
                                                            uint32_t* __restrict__ out, int ntt, Tabs T) {
    __shared__ uint32_t s[kPadN];
    const int c = blockIdx.x, ab = blockIdx.y, p = blockIdx.z;
    const uint32_t q = T.p[p].q;
    const uint64_t* src = in + ((size_t)c * 2 + ab) * kN;
    uint32_t* dst = out + (((size_t)c * 2 + ab) * 2 + p) * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) s[lds_pad(x)] = (uint32_t)(src[x] % q);
    __syncthreads();
    if (ntt) ntt_fwd_r16(s, T.p[p]);
    for (int x = threadIdx.x; x < kN; x += kNttThreads) dst[x] = s[lds_pad(x)];
}

// reorientCipher's packed buffer (twosteps.cpp:1386-1403) from the rotation
// residues: out[z][k][ab] = (x mod q1) | (x mod q2) << 32
__global__ void __launch_bounds__(256) k_ctw_pack(const uint32_t* __restrict__ ctw, uint64_t* __restrict__ out,
                                                  int n1) {
    const size_t total = (size_t)kN * n1 * 2;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int ab = (int)(idx & 1);
        const size_t k = (idx >> 1) % n1;
        const size_t z = (idx >> 1) / n1;
        const uint32_t* w = ctw + (k * 2 + ab) * 2 * kN + z;
        out[idx] = (uint64_t)w[0] | ((uint64_t)w[kN] << 32);
    }
}

// rotation-0 words of the rotation buffer -> u64 [2][N] mod Q
__global__ void __launch_bounds__(256) k_ctw0_to_u64(const uint32_t* __restrict__ ctw, uint64_t* __restrict__ out,
                                                     int n1, Tabs T) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 2 * kN) return;
    const int ab = idx / kN, x = idx % kN;
    const uint32_t* w = ctw + (size_t)(0 * 2 + ab) * 2 * kN + x;
    out[idx] = crt_q(w[0], w[kN], T);
}

// RLWE words mod Q, RNS planes [cnt][2 ab][2 p][N] -> u64 [cnt][2][N]
__global__ void __launch_bounds__(256) k_rns_to_u64(const uint32_t* __restrict__ in, uint64_t* __restrict__ out,
                                                    int cnt, Tabs T) {
    const size_t total = (size_t)cnt * 2 * kN;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const size_t ca = idx / kN;
        const int x = (int)(idx % kN);
        out[idx] = crt_q(in[(ca * 2 + 0) * kN + x], in[(ca * 2 + 1) * kN + x], T);
    }
}

// a[i] += b[i] over RNS planes [cnt][2 ab][2 p][N]
__global__ void __launch_bounds__(256) k_add_planes(uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                                    int cnt, Tabs T) {
    const size_t total = (size_t)cnt * 4 * kN;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int p = (int)((idx / kN) & 1);
        a[idx] = add_mod(a[idx], b[idx], T.p[p].q);
    }
}

// Client recover, decrypt_bsgs_total (twosteps.cpp:388-418 with
// encrypt.cpp:528-573 decrypt): m = (b - a*s) * r^-1 mod Q, rounded to
// round(m / Delta) (m > Q/2: round((m - Q) / Delta) + p).  One workgroup of
// 512 threads per response; each 256-thread half works in one prime.
// secret, response: coefficient form mod Q (u64); msg: N values.
// qinv_r: (r^-1 mod q_p, Shoup companion) for p = 0, 1.
__global__ void __launch_bounds__(512) k_decrypt(const uint64_t* __restrict__ secret, const uint64_t* __restrict__ resp,
                                                 uint64_t* __restrict__ msg, uint2 qinv0, uint2 qinv1, Tabs T) {
    __shared__ uint32_t s[6][kPadN];     // per prime: a, b, s (planes 3p .. 3p+2)
    const int tid = threadIdx.x;
    const int half = __builtin_amdgcn_readfirstlane(tid >> 8), ht = tid & 255;
    const uint64_t* a = resp + (size_t)blockIdx.x * 2 * kN;
    const uint64_t* b = a + kN;
    const uint64_t* sk = secret + (size_t)blockIdx.x * kN;
    uint64_t* out = msg + (size_t)blockIdx.x * kN;
    const PrimeTab& P = T.p[half];
    uint32_t* sa = s[3 * half];
    uint32_t* sb = s[3 * half + 1];
    uint32_t* ss = s[3 * half + 2];
    for (int x = ht; x < kN; x += kNttThreads) {
        const uint64_t va = a[x], vb = b[x], vs = sk[x];
        sa[lds_pad(x)] = half ? (uint32_t)(va % kQ2) : (uint32_t)(va % kQ1);
        sb[lds_pad(x)] = half ? (uint32_t)(vb % kQ2) : (uint32_t)(vb % kQ1);
        ss[lds_pad(x)] = half ? (uint32_t)(vs % kQ2) : (uint32_t)(vs % kQ1);
    }
    __syncthreads();
    ntt_fwd_r16x2_same(sa, sb, P, ht);
    ntt_fwd_r16(ss, P, ht);
    const uint2 qi = half ? qinv1 : qinv0;
    for (int x = ht; x < kN; x += kNttThreads) {
        const uint32_t as = half ? (uint32_t)(((uint64_t)sa[lds_pad(x)] * ss[lds_pad(x)]) % kQ2)
                                 : (uint32_t)(((uint64_t)sa[lds_pad(x)] * ss[lds_pad(x)]) % kQ1);
        const uint32_t m = sub_mod(sb[lds_pad(x)], as, P.q);
        sa[lds_pad(x)] = mul_shoup(m, qi.x, qi.y, P.q);      // (b - a s) * r^-1 mod q_p, NTT form
    }
    __syncthreads();
    ntt_inv_r16(sa, P, ht);
    constexpr uint64_t kDelta = 1021937069741ULL;   // params.h bsgsDelta
    for (int x = tid; x < kN; x += 512) {
        const uint64_t me = crt_q(s[0][lds_pad(x)], s[3][lds_pad(x)], T);
        // round half away from zero of me / Delta, exactly (long double holds
        // these quotients without a rounding that could cross a half)
        uint64_t v;
        if (me > kQ / 2) v = (uint64_t)kP - (2 * (kQ - me) + kDelta) / (2 * kDelta);
        else v = (2 * me + kDelta) / (2 * kDelta);
        out[x] = v;
    }
}

// plain copy, 16 bytes per lane per step (reference kernel for HBM counters)
__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// inverse_encode (twosteps.cpp:182-198) of one record column per workgroup:
// forward NTT mod p (HEXL NTT(N, p) with its minimal root), then the slot
// order of compute_query_decode: msg[dec[j]] = NTT(msg)[j].  msg: [count][N]
// values < p, rewritten in place.
__global__ void __launch_bounds__(kNttThreads) k_inverse_encode(uint64_t* __restrict__ msg,
                                                                const uint16_t* __restrict__ dec, Tabs T) {
    __shared__ uint32_t s[kPadN];
    uint64_t* m = msg + (size_t)blockIdx.x * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) s[lds_pad(x)] = (uint32_t)m[x];
    __syncthreads();
    ntt_fwd_r16(s, T.p[3]);
    for (int x = threadIdx.x; x < kN; x += kNttThreads) m[dec[x]] = s[lds_pad(x)];
}

// test hooks: batched forward / inverse NTT of polynomials in prime p
__global__ void __launch_bounds__(kNttThreads) k_ntt_test(uint32_t* __restrict__ data, int p, int inverse, Tabs T) {
    __shared__ uint32_t s[kPadN];
    uint32_t* d = data + (size_t)blockIdx.x * kN;
    for (int x = threadIdx.x; x < kN; x += kNttThreads) s[lds_pad(x)] = d[x];
    __syncthreads();
    if (inverse) ntt_inv_r16(s, T.p[p]);
    else ntt_fwd_r16(s, T.p[p]);
    for (int x = threadIdx.x; x < kN; x += kNttThreads) d[x] = s[lds_pad(x)];
}

}  // namespace kspir
