// ref_shim.cpp — extern "C" entry points over the reference's own crt.cpp
// (compiled from /root/reference/src by oracle/Makefile into oracle/_ref/).
// TEST INFRASTRUCTURE ONLY: lets the CPU tests cross-check the oracle's
// restatement of the database product and CRT helpers against the
// reference's AVX-512 code on the same inputs.
#include <stdint.h>
#include <vector>
#include "crt.h"

extern "C" {

// fastMultiplyQueryByDatabaseDim1InvCRT (crt.cpp:643): out [n2][2N]
int ref_db_product(uint64_t* out, const uint64_t* db, const uint64_t* cbuf, int n1, int n2) {
    std::vector<std::vector<uint64_t> > o(n2, std::vector<uint64_t>(2 * N, 0));
    fastMultiplyQueryByDatabaseDim1InvCRT(o, db, cbuf, (size_t)n1, (size_t)n2);
    for (int i = 0; i < n2; ++i)
        for (int j = 0; j < 2 * N; ++j) out[(size_t)i * 2 * N + j] = o[i][j];
    return 0;
}

// crt_compose (crt.cpp:368)
uint64_t ref_crt_compose(uint64_t x, uint64_t y) { return crt_compose(x, y); }

// database_tocrt (crt.cpp:859): data_ntt [N/2][N] -> datacrt [N][N2][N1]
int ref_database_tocrt(uint64_t* datacrt, const uint64_t* data_ntt, int n1) {
    std::vector<std::vector<uint64_t> > d(N / 2, std::vector<uint64_t>(N));
    for (int i = 0; i < N / 2; ++i)
        for (int j = 0; j < N; ++j) d[i][j] = data_ntt[(size_t)i * N + j];
    database_tocrt(datacrt, d, n1);
    return 0;
}

}
