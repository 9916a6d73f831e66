// kspir_pir.h — drop-in C++ surface of the KsPIR online server answer on
// MI355X.  It keeps the reference's class and function names, argument
// meanings and visible side effects for the three calls of the online answer
// (tests/test-pir.cpp:111-115):
//
//   matrix_vector_mul_bsgs_rns_crt_large   src/twosteps.h:55
//   externalProduct                        src/lwe.h:292
//   packingRLWEs                           src/answer.h:79
//
// and the server preprocessing of the database (twosteps.h:19
// database_tobsgsntt on the GPU, crt.h:133 database_tocrt), the client side
// of test-pir.cpp:20-24,68-89 (Secret, query_bsgs_rns, RGSW / BSGS /
// packing keyGen: rand() drawn on the host in the reference's order, the
// ring arithmetic on the GPU) and recover (decrypt_bsgs_total)
//
// and the data types they take (src/lwe.h:48-110 RlweCiphertext, :113-169
// RGSWCiphertext, :175-234 AutoKey, :381-466 AutoKeyBSGSRNS), whose public
// members (a, b, keyMap) hold the reference's words.  The work runs on the
// GPU through the C ABI of include/kspir_amd.h; databases and keys are
// uploaded to HBM on first use and kept there (see kspir_pir_forget).
//
// Errors: a missing key index throws std::out_of_range (as keyMap.at does in
// the reference); shapes the reference asserts on, and device failures,
// throw std::runtime_error.
#ifndef KSPIR_PIR_H
#define KSPIR_PIR_H

#include <stdint.h>
#include <iostream>
#include <map>
#include <string>
#include <vector>

namespace kspir_params {
constexpr int32_t N = 4096;                                    // params.h:16
constexpr uint64_t crtq1 = 268369921ULL, crtq2 = 249561089ULL;  // crt.h:16-17
constexpr uint64_t crtMod = crtq1 * crtq2;                       // crt.h:18
constexpr uint64_t bsMod = 16760833ULL;                          // crt.h:32
}  // namespace kspir_params

// src/secret.h:16-61: the RLWE secret.  Secret(module, ntt) draws N
// values rand() % 3 - 1 exactly as secret.cpp:13-60 does (so the caller's
// rand() stream advances as with the reference) and keeps them mod module,
// in NTT form when ntt is set.  Moduli crtMod and bsMod are supported (the
// reference's default bigMod is not used on the test-pir path and throws).
class Secret {
  private:
    uint64_t length = kspir_params::N;
    uint64_t modulus;
    std::vector<uint64_t> coeff;   // coefficient form, the source of every conversion
    std::vector<uint64_t> data;    // current form
    bool nttform;

  public:
    explicit Secret(uint64_t module, bool ntt = true);
    std::vector<uint64_t> getData() { return data; }
    uint64_t getData(int32_t index) { return data.at(index); }
    uint64_t getModulus() { return modulus; }
    int32_t getLength() { return (int32_t)length; }
    bool isNttForm() { return nttform; }
    void toCoeffForm();
    void toNttForm();
    const std::vector<uint64_t>& coefficients() const { return coeff; }
};

// src/lwe.h:48-110
class RlweCiphertext {
  private:
    uint64_t length = kspir_params::N;
    uint64_t modulus = kspir_params::crtMod;
    bool isntt = false;

  public:
    std::vector<uint64_t> a;
    std::vector<uint64_t> b;

    RlweCiphertext();
    RlweCiphertext(uint64_t len, uint64_t module);
    uint64_t getModulus() const { return modulus; }
    void setModulus(uint64_t module) { modulus = module; }
    int32_t getLength() const { return (int32_t)length; }
    bool getIsNtt() const { return isntt; }
    void setIsNtt(bool is) { isntt = is; }
};

// src/lwe.h:113-169 (rows a[0..2ell), b[0..2ell), NTT form)
class RGSWCiphertext {
  private:
    uint64_t length = kspir_params::N;
    uint64_t modulus = kspir_params::crtMod;
    uint64_t ellnum = 2, PP = 1ULL << 20, BBg = 1ULL << 18;
    bool isntt = true;

  public:
    std::vector<std::vector<uint64_t> > a;
    std::vector<std::vector<uint64_t> > b;

    RGSWCiphertext(uint64_t len, uint64_t module, uint64_t elln, uint64_t base, uint64_t bg);
    uint64_t getModulus() const { return modulus; }
    uint64_t getLength() const { return length; }
    uint64_t getEllnum() const { return ellnum; }
    uint64_t getBg() const { return BBg; }
    uint64_t getBase() const { return PP; }
    bool getIsNtt() const { return isntt; }
    void setIsNtt(bool is) { isntt = is; }
    // lwe.h:141, lwe.cpp:297-352: RGSW(X^-index) (is_reverse) or RGSW(X^index),
    // 12N rand() draws; like the reference it leaves the secret in NTT form
    void keyGen(Secret& secret, const uint64_t index, bool is_reverse = false);
};

// src/lwe.h:175-234 (packing keys; keyMap[N/i + 1] = ell RLWE, NTT form)
class AutoKey {
  private:
    uint64_t length = kspir_params::N;
    uint64_t modulus = kspir_params::crtMod;
    int32_t ellnum = 4;
    uint64_t PP = 0, BBg = 1ULL << 14;

  public:
    std::map<int32_t, std::vector<RlweCiphertext> > keyMap;

    AutoKey(int32_t len, uint64_t module, int32_t elln, uint64_t base, uint64_t bg);
    // lwe.h:213, lwe.cpp:509-550: keys for N/i+1, i = 1, 2, .. < num
    // (packing_rlwe) or i+1, i = 2, 4, .. <= num+1, 12N rand() draws each;
    // keyMap.insert keeps existing entries as the reference's does
    void keyGen(Secret& secret, const int32_t num, bool packing_rlwe = false);
    uint64_t getModulus() const { return modulus; }
    uint64_t getLength() const { return length; }
    uint64_t getEllnum() const { return (uint64_t)ellnum; }
    uint64_t getBg() const { return BBg; }
    uint64_t getBase() const { return PP; }
};

// src/lwe.h:381-466 (BSGS keys; keyMap[5^i], i <= N1/2: 3 RLWE mod crtMod
// then 3 mod bsMod; keyMap[5^(N1*i)], i < N2: 2 RLWE mod crtMod)
class AutoKeyBSGSRNS {
  private:
    uint64_t length = kspir_params::N;
    uint64_t modulus = kspir_params::crtMod;
    uint64_t bsModulus = kspir_params::bsMod;

  public:
    std::map<int32_t, std::vector<RlweCiphertext> > keyMap;

    AutoKeyBSGSRNS(int32_t len, uint64_t module, uint64_t bsModule);
    // lwe.h:434, lwe.cpp:1143-1160: baby keys 5^i (i <= N1/2), giant keys
    // 5^(N1*i) (i < N2), in the reference's rand() order
    void bsgsKeyGen(Secret& secret, int32_t N1);
    uint64_t getModulus() const { return modulus; }
    uint64_t getBSModulus() const { return bsModulus; }
    uint64_t getLength() const { return length; }
    uint64_t getEllnumBS() const { return 3; }
    uint64_t getBaseBS() const { return 1ULL << 20; }
    uint64_t getBgBS() const { return 1ULL << 20; }
    uint64_t getEllnumGS() const { return 2; }
    uint64_t getBaseGS() const { return 1ULL << 20; }
    uint64_t getBgGS() const { return 1ULL << 18; }
};

// twosteps.h:55.  input: [0] mod crtMod, [1] mod bsMod, coefficient form;
// as in the reference input[0] becomes the modulus-switched query in NTT
// form.  result: r entries, NTT form.  data: database_tocrt words, r blocks.
void matrix_vector_mul_bsgs_rns_crt_large(std::vector<RlweCiphertext>& result, std::vector<RlweCiphertext>& input,
                                          uint64_t* data, const AutoKeyBSGSRNS& autoKey, const int32_t N1,
                                          std::vector<std::vector<int32_t> >& permutations, int32_t r = 1);

// lwe.h:292.  Accumulates into result (NTT-domain words) and leaves it in
// coefficient form; firstDim is turned into coefficient form if it was NTT.
void externalProduct(RlweCiphertext& result, RlweCiphertext& firstDim, const RGSWCiphertext& query);

// answer.h:79.  rlwes (a power of two of them, coefficient or NTT form) are
// consumed; result is the packed RLWE in coefficient form.
void packingRLWEs(RlweCiphertext& result, std::vector<RlweCiphertext>& rlwes, const AutoKey& autokey);

// twosteps.h:19-21 database_tobsgsntt: data [N][N/2] record values < p;
// result [N/2][N] (pre-sized by the caller) gets the rows mod crtMod in NTT
// form.  modulus is the modulus the centred records are lifted into before
// the transform (twosteps.cpp:155-166): crtMod, as test-pir.cpp:61 passes it
// (the reference transforms mod crtMod whatever modulus says, so any other
// value is rejected).  Runs on the GPU.
void database_tobsgsntt(std::vector<std::vector<uint64_t> >& result, std::vector<std::vector<uint64_t> >& data,
                        uint64_t modulus, int32_t N1);

// crt.h:133 database_tocrt: pack (x mod crtq1) | (x mod crtq2) << 32 into
// datacrt[z][N2][N1] (a re-layout; host).
void database_tocrt(uint64_t* datacrt, std::vector<std::vector<uint64_t> >& data_ntt, int32_t N1);

// twosteps.h:11, twosteps.cpp:23-39 sample_database_bsgs: data (N rows of at
// least N/2 entries) gets rand() & 0xffff per record, drawn on the host in the
// reference's order (row-major), redrawn above p as the reference does (a
// 16-bit draw never is)
void sample_database_bsgs(std::vector<std::vector<uint64_t> >& data);

// twosteps.h:22, twosteps.cpp:182-198 inverse_encode: message (N values < p)
// becomes its forward NTT mod p in the slot order of compute_query_decode.
// Runs on the GPU.
void inverse_encode(std::vector<uint64_t>& message);

// utils.h:129, utils.cpp:1591-1619 compute_permutation_matrix: rows
// 0..max_indexs-1 of permutations (each resized to length) get the NTT-slot
// permutation of tau_{5^i}; length must be N, max_indexs in [1, N/2].  Host,
// closed form.
void compute_permutation_matrix(std::vector<std::vector<int32_t> >& permutations, const int32_t max_indexs,
                                const int32_t length);

// utils.h:21-23, utils.cpp:40-69 showLargeVector: "[v0, v1, .., v15, ..., vlast]"
// (every entry when there are at most 16) on std::cout, after the prefix ss
template <typename T>
void showLargeVector(std::vector<T>& vals) {
    std::cout << "[" << vals[0];
    if (vals.size() <= 16) {
        for (size_t i = 1; i < vals.size(); ++i) std::cout << ", " << vals[i];
    } else {
        for (size_t i = 1; i < 16; ++i) std::cout << ", " << vals[i];
        std::cout << ", ..., " << vals[vals.size() - 1];
    }
    std::cout << "]" << std::endl;
}
void showLargeVector(std::vector<uint64_t>& vals, std::string ss = "");

// twosteps.h:26, twosteps.cpp:279-318: cipher[0] (mod crtMod) and cipher[1]
// (mod bsMod) receive the query for `row`, coefficient form; 5N rand() draws
void query_bsgs_rns(std::vector<RlweCiphertext>& cipher, Secret& queryKey, uint64_t row);

// twosteps.h:30, twosteps.cpp:388-418: message (resized to N) gets the
// recovered plaintext; like the reference, cipher is left in NTT form
void decrypt_bsgs_total(std::vector<uint64_t>& message, RlweCiphertext& cipher, Secret& secret,
                        const int32_t rlwesNum);

// Device copies.  A database pointer or key object is uploaded to HBM on
// first use and reused while its words hash to the same value: every word is
// hashed on every call, so a database or key edited in place is re-uploaded
// and never answered from stale device words.
//   kspir_pir_forget(p)            drops the device copies made for p
//   kspir_pir_assume_unchanged(p)  the caller promises not to change the
//                                  database at p until kspir_pir_forget(p):
//                                  later calls skip hashing its words
void kspir_pir_forget(const void* object_or_data);
void kspir_pir_assume_unchanged(const void* data);

#endif
