// CPU emulation of the register-blocked NTT passes of kspir_device.h
// (TEST INFRASTRUCTURE): runs each pass for all 256 threads in turn, which
// is what the barriers between passes guarantee on the GPU, so the pass
// indexing and twiddle selection can be checked against the oracle on a
// machine without a GPU.
#include <cstdint>
#include <vector>
#include "../../kspir_amd/csrc/kspir_device.h"

using namespace kspir;

namespace {
uint64_t powmod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1;
    b %= q;
    while (e) { if (e & 1) r = (unsigned __int128)r * b % q; b = (unsigned __int128)b * b % q; e >>= 1; }
    return r;
}
uint32_t brev(uint32_t x) { uint32_t r = 0; for (int i = 0; i < kLogN; ++i) { r = (r << 1) | (x & 1); x >>= 1; } return r; }
}

extern "C" int emul_ntt(uint32_t* poly, uint32_t q, uint32_t root, int inverse) {
    std::vector<uint32_t> fw(kN), fwsh(kN), iv(kN), ivsh(kN);
    uint64_t rinv = powmod(root, q - 2, q), x = 1, y = 1;
    for (int i = 0; i < kN; ++i) {
        fw[brev(i)] = (uint32_t)x; iv[brev(i)] = (uint32_t)y;
        x = x * root % q; y = y * rinv % q;
    }
    for (int i = 0; i < kN; ++i) {
        fwsh[i] = (uint32_t)(((uint64_t)fw[i] << 32) / q);
        ivsh[i] = (uint32_t)(((uint64_t)iv[i] << 32) / q);
    }
    const uint32_t ninv = (uint32_t)powmod(kN, q - 2, q), ninvsh = (uint32_t)(((uint64_t)ninv << 32) / q);
    std::vector<uint2> fw2(kN), iv2(kN);
    for (int i = 0; i < kN; ++i) { fw2[i] = make_uint2(0u - fw[i], fwsh[i]); iv2[i] = make_uint2(iv[i], ivsh[i]); }
    uint32_t w[15], ws[15];
    std::vector<uint32_t> s(kPadN, 0);
    for (int e = 0; e < kN; ++e) s[lds_pad(e)] = poly[e];
    uint32_t v[16];
    for (int step = 0; step < 3; ++step) {
        const int pass = inverse ? 2 - step : step;
        for (int tid = 0; tid < 256; ++tid) {
            for (int k = 0; k < 16; ++k) v[k] = s[lds_pad(pass_elem(pass, tid, k))];
            if (inverse) {
                load_pass_twiddles(iv2.data(), pass, tid, w, ws);
                ntt_inv_pass(v, w, ws, q);
                if (pass == 0) ntt_inv_finish(v, ninv, ninvsh, q);
                for (int k = 0; k < 16; ++k)      // lazy words between passes stay below 2q
                    if (v[k] >= (pass == 0 ? q : 2 * q)) return -1;
            } else {
                load_pass_twiddles(fw2.data(), pass, tid, w, ws);
                ntt_fwd_pass(v, w, ws, q);
                if (pass == 2) ntt_fwd_finish(v, q);
                for (int k = 0; k < 16; ++k)      // and below 4q in the forward direction
                    if (v[k] >= (pass == 2 ? q : 4 * q)) return -1;
            }
            for (int k = 0; k < 16; ++k) s[lds_pad(pass_elem(pass, tid, k))] = v[k];
        }
    }
    for (int e = 0; e < kN; ++e) poly[e] = s[lds_pad(e)];
    return 0;
}
