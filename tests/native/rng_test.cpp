// rng_test.cpp — checks kspir_rng.h (seeding, companion matrix, jump-ahead)
// against the C library's own rand() for several seeds and offsets.
// Prints "ok" or the first mismatch.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "kspir_rng.h"

int main(int argc, char** argv) {
    if (argc == 3) {   // rng_test <seed> <k>: output k of the stream after srand(seed)
        uint32_t w[kspir_rng::kLag];
        kspir_rng::seed_window((uint32_t)std::strtoul(argv[1], nullptr, 10), w);
        uint32_t j[kspir_rng::kLag], v;
        kspir_rng::apply(kspir_rng::power(std::strtoull(argv[2], nullptr, 10)), w, j);
        kspir_rng::generate(j, &v, 1);
        std::printf("%u\n", v);
        return 0;
    }
    const unsigned seeds[] = {1, 2, 7, 12345, 0xdeadbeefu, 0};
    for (unsigned seed : seeds) {
        srand(seed);
        std::vector<uint32_t> ref(200000);
        for (auto& v : ref) v = (uint32_t)rand();
        uint32_t w[kspir_rng::kLag];
        kspir_rng::seed_window(seed, w);
        std::vector<uint32_t> seq(ref.size());
        uint32_t w2[kspir_rng::kLag];
        std::memcpy(w2, w, sizeof w);
        kspir_rng::generate(w2, seq.data(), seq.size());
        for (size_t k = 0; k < ref.size(); ++k)
            if (seq[k] != ref[k]) { std::printf("seed %u: sequential output %zu differs\n", seed, k); return 1; }
        for (uint64_t off : {0ull, 1ull, 30ull, 31ull, 4096ull, 65537ull, 123457ull, 199000ull}) {
            uint32_t j[kspir_rng::kLag];
            kspir_rng::apply(kspir_rng::power(off), w, j);
            uint32_t got[64];
            kspir_rng::generate(j, got, 64);
            for (int k = 0; k < 64 && off + k < ref.size(); ++k)
                if (got[k] != ref[off + k]) {
                    std::printf("seed %u: jump %llu output %d differs\n", seed, (unsigned long long)off, k);
                    return 1;
                }
        }
    }
    std::printf("ok\n");
    return 0;
}
