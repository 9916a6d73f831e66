// dropin_test.cpp — the online answer of tests/test-pir.cpp:107-116, written
// against the drop-in C++ surface (kspir_amd/cpp/kspir_pir.h) exactly as the
// reference driver calls it.  TEST INFRASTRUCTURE: inputs come from raw files
// written by the test (oracle-generated, reference layouts); outputs go back
// as raw files for word-for-word comparison.
//   dropin_test <dir> <r> <N1>  |  client ...  |  decrypt ...  |  sample <dir> <r> <seed>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "kspir_pir.h"

using namespace kspir_params;

static std::vector<uint64_t> read_u64(const std::string& path, size_t n) {
    std::vector<uint64_t> v(n);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), 8, n, f) != n) { std::fprintf(stderr, "cannot read %s\n", path.c_str()); std::exit(2); }
    std::fclose(f);
    return v;
}
static std::vector<int32_t> read_i32(const std::string& path, size_t n) {
    std::vector<int32_t> v(n);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), 4, n, f) != n) { std::fprintf(stderr, "cannot read %s\n", path.c_str()); std::exit(2); }
    std::fclose(f);
    return v;
}
static void write_u64(const std::string& path, const std::vector<uint64_t>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    std::fwrite(v.data(), 8, v.size(), f);
    std::fclose(f);
}
static RlweCiphertext rlwe(const uint64_t* p, uint64_t q, bool ntt) {
    RlweCiphertext c(N, q);
    c.a.assign(p, p + N);
    c.b.assign(p + N, p + 2 * N);
    c.setIsNtt(ntt);
    return c;
}
static uint32_t pow5(int e) { uint32_t r = 1; for (int i = 0; i < e; ++i) r = (r * 5u) & (2 * N - 1); return r; }

// test-pir.cpp:20-24,68-89 through the drop-in: srand(seed), Secret,
// query_bsgs_rns, RGSW keyGen (is_reverse), bsgsKeyGen, packing keyGen; the
// words go to <dir> in the layouts of include/kspir_amd.h
static int client_mode(const std::string& dir, int32_t r, int32_t N1, unsigned seed, int tcol, int tpack) {
    srand(seed);
    Secret queryKey(crtMod, false);
    std::vector<RlweCiphertext> query1(1, RlweCiphertext(N, crtMod));
    query1.push_back(RlweCiphertext(N, bsMod));
    query_bsgs_rns(query1, queryKey, (uint64_t)tcol);
    RGSWCiphertext queryGsw(N, crtMod, 2, 1 << 20, 1 << 18);
    queryGsw.keyGen(queryKey, (uint64_t)tpack, true);
    AutoKeyBSGSRNS autoKey(N, crtMod, bsMod);
    autoKey.bsgsKeyGen(queryKey, N1);
    AutoKey packingKey(N, crtMod, 4, 0, 1 << 14);
    packingKey.keyGen(queryKey, r, true);
    const int N2 = N / 2 / N1;
    std::vector<uint64_t> q, g, bb, gg, pk;
    for (int i = 0; i < 2; ++i) {
        q.insert(q.end(), query1[i].a.begin(), query1[i].a.end());
        q.insert(q.end(), query1[i].b.begin(), query1[i].b.end());
    }
    for (int k = 0; k < 4; ++k) {
        g.insert(g.end(), queryGsw.a[k].begin(), queryGsw.a[k].end());
        g.insert(g.end(), queryGsw.b[k].begin(), queryGsw.b[k].end());
    }
    for (int i = 1; i <= N1 / 2; ++i)
        for (const RlweCiphertext& c : autoKey.keyMap.at(pow5(i))) {
            bb.insert(bb.end(), c.a.begin(), c.a.end());
            bb.insert(bb.end(), c.b.begin(), c.b.end());
        }
    for (int i = 1; i < N2; ++i)
        for (const RlweCiphertext& c : autoKey.keyMap.at(pow5(N1 * i))) {
            gg.insert(gg.end(), c.a.begin(), c.a.end());
            gg.insert(gg.end(), c.b.begin(), c.b.end());
        }
    for (int i = 1; i < r; i <<= 1)
        for (const RlweCiphertext& c : packingKey.keyMap.at(N / i + 1)) {
            pk.insert(pk.end(), c.a.begin(), c.a.end());
            pk.insert(pk.end(), c.b.begin(), c.b.end());
        }
    const int ntt_after = (int)queryKey.isNttForm();
    queryKey.toCoeffForm();
    write_u64(dir + "/secret.bin", queryKey.getData());
    write_u64(dir + "/query.bin", q);
    write_u64(dir + "/rgsw.bin", g);
    write_u64(dir + "/baby.bin", bb);
    write_u64(dir + "/giant.bin", gg);
    write_u64(dir + "/packing.bin", pk);
    std::printf("client seed %u: secret in ntt form after the RGSW keyGen: %d\n", seed, ntt_after);
    return 0;
}

// recover through the drop-in: srand(seed), Secret, decrypt_bsgs_total of
// <dir>/response.bin -> message.bin and the mutated response (NTT form)
static int decrypt_mode(const std::string& dir, int32_t r, unsigned seed) {
    srand(seed);
    Secret queryKey(crtMod, false);
    std::vector<uint64_t> rw = read_u64(dir + "/response.bin", 2 * (size_t)N);
    RlweCiphertext resp = rlwe(rw.data(), crtMod, false);
    std::vector<uint64_t> msg;
    decrypt_bsgs_total(msg, resp, queryKey, r);
    write_u64(dir + "/message.bin", msg);
    std::vector<uint64_t> after(resp.a);
    after.insert(after.end(), resp.b.begin(), resp.b.end());
    write_u64(dir + "/response_after.bin", after);
    std::printf("decrypt: response ntt=%d\n", (int)resp.getIsNtt());
    return 0;
}

// sample_database_bsgs through the drop-in: srand(seed), r blocks of
// records drawn in test-pir's order -> <dir>/records.bin ([r][N][N/2])
static int sample_mode(const std::string& dir, int32_t r, unsigned seed) {
    srand(seed);
    std::vector<std::vector<uint64_t> > data(N, std::vector<uint64_t>(N / 2, 0));
    std::vector<uint64_t> flat;
    for (int k = 0; k < r; ++k) {
        sample_database_bsgs(data);
        for (int i = 0; i < N; ++i) flat.insert(flat.end(), data[i].begin(), data[i].end());
    }
    write_u64(dir + "/records.bin", flat);
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 5 && std::string(argv[1]) == "sample")
        return sample_mode(argv[2], std::atoi(argv[3]), (unsigned)std::strtoul(argv[4], nullptr, 10));
    if (argc == 8 && std::string(argv[1]) == "client")
        return client_mode(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), (unsigned)std::strtoul(argv[5], nullptr, 10),
                           std::atoi(argv[6]), std::atoi(argv[7]));
    if (argc == 5 && std::string(argv[1]) == "decrypt")
        return decrypt_mode(argv[2], std::atoi(argv[3]), (unsigned)std::strtoul(argv[4], nullptr, 10));
    if (argc != 4) { std::fprintf(stderr, "usage: dropin_test <dir> <r> <N1> | client <dir> <r> <N1> <seed> <col> <pack> | decrypt <dir> <r> <seed>\n"); return 2; }
    const std::string dir = argv[1];
    const int32_t r = std::atoi(argv[2]), N1 = std::atoi(argv[3]), N2 = N / 2 / N1;
    int lr = 0;
    while ((1 << lr) < r) ++lr;
    std::vector<uint64_t> datacrt = read_u64(dir + "/db.bin", (size_t)r * N * N / 2);
    // server preprocessing as test-pir.cpp:58-62 runs it, when records are given
    if (FILE* f = std::fopen((dir + "/records.bin").c_str(), "rb")) {
        std::fclose(f);
        const size_t words = (size_t)N * N / 2;
        std::vector<uint64_t> rec = read_u64(dir + "/records.bin", (size_t)r * words), pre((size_t)r * words);
        std::vector<std::vector<uint64_t> > data(N, std::vector<uint64_t>(N / 2)), data_ntt(N / 2, std::vector<uint64_t>(N));
        for (int b = 0; b < r; ++b) {
            for (int i = 0; i < N; ++i) data[i].assign(&rec[b * words + (size_t)i * (N / 2)], &rec[b * words + (size_t)(i + 1) * (N / 2)]);
            database_tobsgsntt(data_ntt, data, kspir_params::crtMod, N1);
            database_tocrt(&pre[b * words], data_ntt, N1);
        }
        write_u64(dir + "/datacrt_gpu.bin", pre);
    }
    std::vector<uint64_t> qw = read_u64(dir + "/query.bin", 4 * (size_t)N);
    std::vector<uint64_t> rg = read_u64(dir + "/rgsw.bin", 8 * (size_t)N);
    std::vector<uint64_t> baby = read_u64(dir + "/baby.bin", (size_t)(N1 / 2) * 12 * N);
    std::vector<uint64_t> giant = read_u64(dir + "/giant.bin", (size_t)(N2 - 1) * 4 * N);
    std::vector<uint64_t> pack = read_u64(dir + "/packing.bin", (size_t)lr * 8 * N);
    std::vector<int32_t> pm = read_i32(dir + "/perms.bin", (size_t)N1 * N);

    // the objects test-pir.cpp:68-98 builds
    std::vector<RlweCiphertext> query1(1, rlwe(&qw[0], crtMod, false));
    query1.push_back(rlwe(&qw[2 * N], bsMod, false));
    RGSWCiphertext queryGsw(N, crtMod, 2, 1 << 20, 1 << 18);
    for (int k = 0; k < 4; ++k) {
        queryGsw.a[k].assign(&rg[(size_t)k * 2 * N], &rg[(size_t)k * 2 * N + N]);
        queryGsw.b[k].assign(&rg[(size_t)k * 2 * N + N], &rg[(size_t)(k + 1) * 2 * N]);
    }
    AutoKeyBSGSRNS autoKey(N, crtMod, bsMod);
    for (int i = 1; i <= N1 / 2; ++i) {
        std::vector<RlweCiphertext> v;
        for (int j = 0; j < 6; ++j) v.push_back(rlwe(&baby[(((size_t)(i - 1) * 6 + j) * 2) * N], j < 3 ? crtMod : bsMod, true));
        autoKey.keyMap[pow5(i)] = v;
    }
    for (int i = 1; i < N2; ++i) {
        std::vector<RlweCiphertext> v;
        for (int j = 0; j < 2; ++j) v.push_back(rlwe(&giant[(((size_t)(i - 1) * 2 + j) * 2) * N], crtMod, true));
        autoKey.keyMap[pow5(N1 * i)] = v;
    }
    AutoKey packingKey(N, crtMod, 4, 0, 1 << 14);
    for (int i = 1, li = 0; i < r; i <<= 1, ++li) {
        std::vector<RlweCiphertext> v;
        for (int j = 0; j < 4; ++j) v.push_back(rlwe(&pack[(((size_t)li * 4 + j) * 2) * N], crtMod, true));
        packingKey.keyMap[N / i + 1] = v;
    }
    std::vector<std::vector<int32_t> > permutations(N1, std::vector<int32_t>(N));
    for (int i = 0; i < N1; ++i) permutations[i].assign(&pm[(size_t)i * N], &pm[(size_t)(i + 1) * N]);

    std::vector<RlweCiphertext> result(r, RlweCiphertext(N, crtMod));
    RlweCiphertext result_output(N, crtMod);
    std::vector<RlweCiphertext> ext_rlwes;
    for (int i = 0; i < r; ++i) ext_rlwes.push_back(RlweCiphertext(N, crtMod));

    // test-pir.cpp:107-116
    auto start = std::chrono::high_resolution_clock::now();
    matrix_vector_mul_bsgs_rns_crt_large(result, query1, datacrt.data(), autoKey, N1, permutations, r);
    std::vector<uint64_t> fd((size_t)r * 2 * N);
    for (int j = 0; j < r; ++j) {
        std::copy(result[j].a.begin(), result[j].a.end(), fd.begin() + (size_t)j * 2 * N);
        std::copy(result[j].b.begin(), result[j].b.end(), fd.begin() + (size_t)j * 2 * N + N);
    }
    for (int j = 0; j < r; ++j) externalProduct(ext_rlwes[j], result[j], queryGsw);
    packingRLWEs(result_output, ext_rlwes, packingKey);
    auto stop = std::chrono::high_resolution_clock::now();

    std::vector<uint64_t> resp(2 * (size_t)N);
    std::copy(result_output.a.begin(), result_output.a.end(), resp.begin());
    std::copy(result_output.b.begin(), result_output.b.end(), resp.begin() + N);
    write_u64(dir + "/response.bin", resp);
    write_u64(dir + "/firstdim.bin", fd);
    std::vector<uint64_t> q0(2 * (size_t)N);
    std::copy(query1[0].a.begin(), query1[0].a.end(), q0.begin());
    std::copy(query1[0].b.begin(), query1[0].b.end(), q0.begin() + N);
    write_u64(dir + "/query0_after.bin", q0);

    // one database word edited in place, then the first dimension again: the
    // drop-in must answer from the edited words (every word is hashed on every
    // call), never from its device copy of the old ones
    const size_t words = (size_t)N * N / 2;
    const size_t w = (size_t)(r > 1 ? 1 : 0) * words + ((size_t)100 * N2 + (N2 > 2 ? 2 : 0)) * N1 + 5;
    const uint64_t lo = datacrt[w] & 0xffffffffULL, hi = datacrt[w] >> 32;
    datacrt[w] = ((lo + 1) % crtq1) | (hi << 32);
    write_u64(dir + "/edit_index.bin", std::vector<uint64_t>{w, datacrt[w]});
    std::vector<RlweCiphertext> query2(1, rlwe(&qw[0], crtMod, false));
    query2.push_back(rlwe(&qw[2 * N], bsMod, false));
    std::vector<RlweCiphertext> result2(r, RlweCiphertext(N, crtMod));
    matrix_vector_mul_bsgs_rns_crt_large(result2, query2, datacrt.data(), autoKey, N1, permutations, r);
    for (int j = 0; j < r; ++j) {
        std::copy(result2[j].a.begin(), result2[j].a.end(), fd.begin() + (size_t)j * 2 * N);
        std::copy(result2[j].b.begin(), result2[j].b.end(), fd.begin() + (size_t)j * 2 * N + N);
    }
    write_u64(dir + "/firstdim_edit.bin", fd);
    std::printf("online server response costs %lld us (first call, includes uploads); input[0] ntt=%d result ntt=%d\n",
                (long long)std::chrono::duration_cast<std::chrono::microseconds>(stop - start).count(),
                (int)query1[0].getIsNtt(), (int)result_output.getIsNtt());
    return 0;
}
