// sanitize_host.cpp — host code under AddressSanitizer and UndefinedBehavior
// Sanitizer (TEST INFRASTRUCTURE): the CPU oracle, the rand() restatement of
// kspir_rng.h and the host-only paths of the C++ drop-in (kspir_pir.cpp),
// built with -fsanitize=address,undefined and run on small inputs.  Any
// report aborts the program; on success it prints "sanitize ok".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "kspir_oracle.h"
#include "kspir_pir.h"
#include "kspir_rng.h"

static int fails = 0;
#define CHECK(c, msg) do { if (!(c)) { std::printf("FAIL: %s\n", msg); ++fails; } } while (0)

int main() {
    const int N = 4096, r = 2, n1 = 128, n2 = N / 2 / n1;
    const size_t words = (size_t)N * N / 2;
    ko_set_threads(2);

    // oracle: the test-pir flow at r 2, seed 1, recovers [1..N]
    {
        std::vector<uint64_t> secret(N), db(r * words), query(4 * N), rgsw(8 * N), baby((size_t)(n1 / 2) * 12 * N),
            giant((size_t)(n2 - 1) * 4 * N), packing(8 * N), resp(2 * N), msg(N);
        std::vector<int32_t> target(2), perms((size_t)n1 * N);
        CHECK(ko_gen_instance(r, n1, 1, 1, secret.data(), target.data(), db.data(), query.data(), rgsw.data(),
                              baby.data(), giant.data(), packing.data(), perms.data()) == 0, "gen_instance");
        CHECK(ko_answer(r, n1, db.data(), query.data(), rgsw.data(), baby.data(), giant.data(), packing.data(),
                        perms.data(), resp.data(), nullptr, nullptr, nullptr) == 0, "answer");
        CHECK(ko_decrypt(r, secret.data(), resp.data(), msg.data()) == 0, "decrypt");
        bool kat = true;
        for (int i = 0; i < N; ++i) kat = kat && msg[i] == (uint64_t)(i + 1);
        CHECK(kat, "oracle KAT [1..N]");
        // the drop-in's host permutation table equals the oracle's
        std::vector<std::vector<int32_t> > pm(n1, std::vector<int32_t>(N));
        compute_permutation_matrix(pm, n1, N);
        bool same = true;
        for (int i = 0; i < n1; ++i) same = same && std::memcmp(pm[i].data(), &perms[(size_t)i * N], N * 4) == 0;
        CHECK(same, "compute_permutation_matrix");
    }
    // decompositions on boundary words
    {
        const uint64_t Q = ko_const(3);
        std::vector<uint64_t> x(N), out(4 * N), out2(4 * N), xb(N);
        for (int i = 0; i < N; ++i) { x[i] = (uint64_t)i * 0x9E3779B97F4A7ull % (Q + 1); xb[i] = (uint64_t)i * 977 % 16760833; }
        x[0] = 0; x[1] = Q; x[2] = Q - 1;
        ko_decompose(out.data(), x.data(), 4, 0, 1 << 14, Q);
        x[1] = Q - 1;
        ko_decompose_rlwe(out2.data(), x.data(), x.data(), 2, 1 << 20, 1 << 18, Q);
        ko_decompose_bsgs(out.data(), out2.data(), x.data(), xb.data(), 3, 1 << 20, 1 << 20);
    }
    // rand() restatement against the C library
    for (unsigned seed : {1u, 77u, 0xabcdefu}) {
        srand(seed);
        std::vector<uint32_t> ref(50000), got(50000);
        for (auto& v : ref) v = (uint32_t)rand();
        uint32_t w[kspir_rng::kLag];
        kspir_rng::seed_window(seed, w);
        kspir_rng::generate(w, got.data(), got.size());
        CHECK(ref == got, "kspir_rng stream");
        uint32_t w0[kspir_rng::kLag], wj[kspir_rng::kLag], v = 0;
        kspir_rng::seed_window(seed, w0);
        kspir_rng::apply(kspir_rng::power(31337), w0, wj);
        kspir_rng::generate(wj, &v, 1);
        CHECK(v == ref[31337], "kspir_rng jump-ahead");
    }
    // drop-in host paths: record sampling in the reference's order, the
    // datacrt re-layout, printing, and argument checks that throw before any
    // device work
    {
        std::vector<uint64_t> rec(r * words), want(r * words);
        CHECK(ko_gen_records(r, n1, 9, 0, 0, 0, rec.data(), want.data()) == 0, "gen_records");
        srand(9);
        std::vector<std::vector<uint64_t> > data(N, std::vector<uint64_t>(N / 2));
        bool same = true;
        for (int b = 0; b < r; ++b) {
            sample_database_bsgs(data);
            for (int i = 0; i < N && same; ++i) same = std::memcmp(data[i].data(), &rec[b * words + (size_t)i * (N / 2)], (N / 2) * 8) == 0;
        }
        CHECK(same, "sample_database_bsgs");
        std::vector<uint64_t> nt((size_t)(N / 2) * N), crt(words);
        CHECK(ko_database_tobsgsntt(n1, rec.data(), nt.data()) == 0, "database_tobsgsntt");
        std::vector<std::vector<uint64_t> > ntv(N / 2);
        for (int i = 0; i < N / 2; ++i) ntv[i].assign(&nt[(size_t)i * N], &nt[(size_t)(i + 1) * N]);
        database_tocrt(crt.data(), ntv, n1);
        CHECK(std::memcmp(crt.data(), want.data(), words * 8) == 0, "database_tocrt");
        std::vector<uint64_t> v(20);
        for (int i = 0; i < 20; ++i) v[i] = i + 1;
        showLargeVector(v, "show: ");
        std::vector<RlweCiphertext> result(r, RlweCiphertext(N, kspir_params::crtMod)), in(2);
        AutoKeyBSGSRNS ak(N, kspir_params::crtMod, kspir_params::bsMod);
        std::vector<std::vector<int32_t> > pm(n1, std::vector<int32_t>(N));
        bool threw = false;
        try { matrix_vector_mul_bsgs_rns_crt_large(result, in, crt.data(), ak, 96, pm, r); } catch (const std::runtime_error&) { threw = true; }
        CHECK(threw, "N1 = 96 rejected");
        threw = false;
        std::vector<std::vector<int32_t> > small(3);
        try { compute_permutation_matrix(small, 8, N); } catch (const std::runtime_error&) { threw = true; }
        CHECK(threw, "short permutation matrix rejected");
    }
    if (fails) return 1;
    std::printf("sanitize ok\n");
    return 0;
}
