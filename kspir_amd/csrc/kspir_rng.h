// kspir_rng.h — the glibc rand() stream (TYPE_3 additive feedback generator,
// the default state of srand/rand) with jump-ahead, host side.
//
// The generator, as glibc documents it in random_r.c: seeding fills
// r[0] = seed (1 if seed is 0), r[i] = 16807 * r[i-1] mod (2^31 - 1) for
// i = 1..30 (Schrage's method, kept in [0, 2^31 - 1)), r[i] = r[i-31] for
// i = 31..33, then r[i] = r[i-31] + r[i-3] mod 2^32 for i >= 34; the first
// 310 of those are discarded and output k of rand() is r[344 + k] >> 1.
// The recurrence is linear over Z/2^32, so with the 31-word window
// S_t = (r[t-31], ..., r[t-1]) and its companion matrix M, S_{t+n} = M^n S_t.
// Header-only so a CPU test can check it against the C library's rand().
#pragma once
#include <cstdint>
#include <cstring>
#include <vector>

namespace kspir_rng {

constexpr int kLag = 31;

// window S_344 after srand(seed): r[313..343]
inline void seed_window(uint32_t seed, uint32_t out[kLag]) {
    std::vector<uint32_t> r(344);
    int32_t w = seed == 0 ? 1 : (int32_t)seed;
    r[0] = (uint32_t)w;
    for (int i = 1; i < 31; ++i) {
        const int32_t hi = w / 127773, lo = w % 127773;
        w = 16807 * lo - 2836 * hi;
        if (w < 0) w += 2147483647;
        r[i] = (uint32_t)w;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 34; i < 344; ++i) r[i] = r[i - 31] + r[i - 3];
    for (int i = 0; i < kLag; ++i) out[i] = r[313 + i];
}

// row-major 31 x 31 matrices over Z/2^32
using Mat = std::vector<uint32_t>;

inline Mat companion() {
    Mat m(kLag * kLag, 0);
    for (int i = 0; i + 1 < kLag; ++i) m[i * kLag + i + 1] = 1;   // shift
    m[(kLag - 1) * kLag + 0] = 1;                                 // r[t-31]
    m[(kLag - 1) * kLag + kLag - 3] += 1;                         // r[t-3]
    return m;
}

inline Mat mul(const Mat& a, const Mat& b) {
    Mat c(kLag * kLag, 0);
    for (int i = 0; i < kLag; ++i)
        for (int k = 0; k < kLag; ++k) {
            const uint32_t x = a[i * kLag + k];
            if (!x) continue;
            for (int j = 0; j < kLag; ++j) c[i * kLag + j] += x * b[k * kLag + j];
        }
    return c;
}

// M^n by repeated squaring
inline Mat power(uint64_t n) {
    Mat r(kLag * kLag, 0), b = companion();
    for (int i = 0; i < kLag; ++i) r[i * kLag + i] = 1;
    while (n) {
        if (n & 1) r = mul(r, b);
        b = mul(b, b);
        n >>= 1;
    }
    return r;
}

inline void apply(const Mat& m, const uint32_t in[kLag], uint32_t out[kLag]) {
    uint32_t t[kLag];
    for (int i = 0; i < kLag; ++i) {
        uint32_t s = 0;
        for (int j = 0; j < kLag; ++j) s += m[i * kLag + j] * in[j];
        t[i] = s;
    }
    std::memcpy(out, t, sizeof t);
}

// the next `count` outputs from window w (advanced in place)
inline void generate(uint32_t w[kLag], uint32_t* out, size_t count) {
    for (size_t k = 0; k < count; ++k) {
        const uint32_t v = w[0] + w[kLag - 3];
        std::memmove(w, w + 1, (kLag - 1) * sizeof(uint32_t));
        w[kLag - 1] = v;
        out[k] = v >> 1;
    }
}

}  // namespace kspir_rng
