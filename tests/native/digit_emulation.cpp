// CPU calls of the device digit decompositions of kspir_device.h (TEST
// INFRASTRUCTURE): the same __host__ __device__ functions the kernels use,
// run on host inputs so that their boundary cases (d == Q, d == M, x == Q,
// x == 0) can be compared with the oracle without a GPU.  Each entry point
// writes the digits of n inputs as the kernels store them: the signed digit
// taken mod the modulus.
#include <cstdint>
#include "../../kspir_amd/csrc/kspir_device.h"

using namespace kspir;

namespace {
uint64_t to_mod(int32_t d, uint64_t q) { return d < 0 ? q - (uint64_t)(-(int64_t)d) : (uint64_t)d; }
}

// giant steps (k_giant_*): decompose ell 2, PP 2^20, Bg 2^18, inputs in [0, Q]; out [2][n] mod Q
extern "C" void emul_digits_giant(const uint64_t* x, int n, uint64_t* out) {
    for (int j = 0; j < 2; ++j)
        for (int i = 0; i < n; ++i) out[(size_t)j * n + i] = to_mod(tfhe_digit<18, 20, 2>(x[i], j), kQ);
}

// packing (k_pack_*): decompose ell 4, PP 0, Bg 2^14, inputs in [0, Q]; out [4][n] mod Q
extern "C" void emul_digits_packing(const uint64_t* x, int n, uint64_t* out) {
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < n; ++i) out[(size_t)j * n + i] = to_mod(tfhe_digit<14, 0, 4>(x[i], j), kQ);
}

// external product (k_ext_*): decompose_rlwe ell 2, PP 2^20, Bg 2^18, inputs in [0, Q); out [2][n] mod Q
extern "C" void emul_digits_ext(const uint64_t* x, int n, uint64_t* out) {
    for (int j = 0; j < 2; ++j)
        for (int i = 0; i < n; ++i) out[(size_t)j * n + i] = to_mod(signed_digit<18, 20>(x[i], j), kQ);
}

// baby steps (k_baby_prep, k_baby_mid_digits): decompose_bsgs ell 3, PP = Bg
// = 2^20 of the composed value v = lo | hi << 64 in [0, Q * bsMod); out_q,
// out_bs [3][n] mod Q and mod bsMod
extern "C" void emul_digits_bsgs(const uint64_t* lo, const uint64_t* hi, int n, uint64_t* out_q, uint64_t* out_bs) {
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < n; ++i) {
            const unsigned __int128 v = (unsigned __int128)hi[i] << 64 | lo[i];
            const int32_t d = bsgs_digit(v, j);
            out_q[(size_t)j * n + i] = to_mod(d, kQ);
            out_bs[(size_t)j * n + i] = digit_residue(d, kBs);
        }
}
