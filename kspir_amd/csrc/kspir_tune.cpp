// kspir_tune.cpp — the autotuner: measures the answer pipeline for one
// database shape on this device and sets the context options that choose
// kernel shapes and paths.  Written against the public C ABI only
// (kspir_context_set/get and the HIP-event timing entry points), so it
// measures exactly what callers run.  No counterpart in the reference, whose
// shapes are compile-time constants of a CPU loop.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kspir_amd.h"
#include "kspir_amd_runtime.h"

// sets the text kspir_last_error reports (defined beside it, not exported)
extern "C" __attribute__((visibility("hidden"))) int kspir_set_error(int code, const char* msg);

namespace {

struct Buf {
    kspir_context* c;
    void* p = nullptr;
    Buf(kspir_context* ctx, size_t bytes) : c(ctx) {
        if (kspir_device_alloc(c, bytes, &p) != KSPIR_OK) p = nullptr;
    }
    ~Buf() {
        if (p) kspir_device_free(c, p);
    }
};

std::string fmt(const char* f, double v) {
    char b[64];
    std::snprintf(b, sizeof b, f, v);
    return b;
}

}  // namespace

extern "C" int kspir_autotune(kspir_context* c, const kspir_database* db, const kspir_keys* const* keys,
                              int32_t nkeys, int32_t max_batch, double latency_budget_ms, int32_t* serve_batch,
                              char* report, size_t report_len) {
    if (!c || !db || !keys || nkeys < 1 || max_batch < 1)
        return kspir_set_error(KSPIR_EINVAL, "kspir_autotune: null handle, nkeys < 1 or max_batch < 1");
    for (int32_t i = 0; i < nkeys; ++i)
        if (!keys[i]) return kspir_set_error(KSPIR_EINVAL, "kspir_autotune: null key handle");
    // zero words are valid ciphertext words in every modulus
    const size_t qw = (size_t)4 * KSPIR_N, rw = (size_t)2 * KSPIR_N;
    Buf qd(c, (size_t)max_batch * qw * 8), rd(c, (size_t)max_batch * rw * 8);
    if (!qd.p || !rd.p) return kspir_set_error(KSPIR_ENOMEM, "kspir_autotune: no device memory for the timing batch");
    {
        std::vector<uint64_t> zero((size_t)max_batch * qw, 0);
        int rc = kspir_memcpy_h2d(c, qd.p, zero.data(), zero.size() * 8);
        if (rc) return rc;
    }
    std::vector<const kspir_keys*> kl(max_batch);
    for (int i = 0; i < max_batch; ++i) kl[i] = keys[i % nkeys];
    const uint64_t* q = (const uint64_t*)qd.p;
    uint64_t* r = (uint64_t*)rd.p;
    auto step_ms = [&](int b, double* ms) -> int {
        float t = 0;
        const int iters = b >= 16 ? 4 : 10;
        int rc = kspir_time_answer_batch(c, db, kl.data(), b, q, r, 2, iters, &t);
        *ms = t / iters;
        return rc;
    };
    std::string js = "{";
    int rc;

    // 1. workgroup shape of the single-query database pass
    int best_w = 4;
    double best_t = 1e30;
    js += "\"db_waves_ms\": {";
    for (int w : {2, 4, 8, 16}) {
        if ((rc = kspir_context_set(c, "db_waves", w))) return rc;
        float t = 0;
        if ((rc = kspir_time_db_pass(c, db, kl[0], q, 10, &t))) return rc;
        js += (w == 2 ? "" : ", ") + std::string("\"") + std::to_string(w) + "\": " + fmt("%.4f", t);
        if (t < best_t) { best_t = t; best_w = w; }
    }
    js += "}, \"db_waves\": " + std::to_string(best_w);
    if ((rc = kspir_context_set(c, "db_waves", best_w))) return rc;

    std::vector<int> sizes;
    for (int b = 1; b <= max_batch; b *= 2) sizes.push_back(b);
    if (sizes.back() != max_batch) sizes.push_back(max_batch);

    // 2. database pass per batch size: int8-MFMA limb GEMM against the VALU
    // stream; the threshold is the smallest size from which MFMA keeps winning
    int mfma_from = max_batch + 1;
    js += ", \"db_pass_ms\": {";
    std::vector<bool> mfma_wins;
    for (size_t i = 0; i < sizes.size(); ++i) {
        const int b = sizes[i];
        double tv = 0, tm = 0;
        if ((rc = kspir_context_set(c, "mfma_min_batch", b + 1))) return rc;   // VALU
        if ((rc = step_ms(b, &tv))) return rc;
        if (b >= 2) {
            if ((rc = kspir_context_set(c, "mfma_min_batch", b))) return rc;   // MFMA
            if ((rc = step_ms(b, &tm))) return rc;
        } else {
            tm = 1e30;
        }
        mfma_wins.push_back(tm < tv);
        js += (i ? ", " : "") + std::string("\"") + std::to_string(b) + "\": [" + fmt("%.4f", tv) + ", " +
              (b >= 2 ? fmt("%.4f", tm) : std::string("null")) + "]";
    }
    for (int i = (int)sizes.size() - 1; i >= 0 && mfma_wins[i]; --i) mfma_from = sizes[i];
    js += "}, \"mfma_min_batch\": " + std::to_string(mfma_from);
    if ((rc = kspir_context_set(c, "mfma_min_batch", mfma_from))) return rc;

    // 3. queries per MFMA sweep at the largest batch (when it takes MFMA)
    if (max_batch >= mfma_from) {
        int best_chunk = 1 << 30;
        double bt = 1e30;
        js += ", \"mfma_chunk_ms\": {";
        bool first = true;
        for (int ch : {8, 16, 32, 64, 1 << 30}) {
            if (ch != (1 << 30) && ch >= max_batch) continue;
            if ((rc = kspir_context_set(c, "mfma_max_chunk", ch))) return rc;
            double t = 0;
            if ((rc = step_ms(max_batch, &t))) return rc;
            js += (first ? "" : ", ") + std::string("\"") + (ch == (1 << 30) ? std::string("all") : std::to_string(ch)) +
                  "\": " + fmt("%.4f", t);
            first = false;
            if (t < bt) { bt = t; best_chunk = ch; }
        }
        js += "}, \"mfma_max_chunk\": " + std::to_string(best_chunk);
        if ((rc = kspir_context_set(c, "mfma_max_chunk", best_chunk))) return rc;
    }

    // 4. giant steps (fused per (query, block) against spread per (block,
    // row, prime)) and packing (fused level against three kernels)
    int fused_from = max_batch + 1;
    std::vector<bool> fused_wins;
    js += ", \"giant_ms\": {";
    for (size_t i = 0; i < sizes.size(); ++i) {
        const int b = sizes[i];
        double ts = 0, tf = 0;
        if ((rc = kspir_context_set(c, "fused_min_batch", b + 1))) return rc;
        if ((rc = step_ms(b, &ts))) return rc;
        if ((rc = kspir_context_set(c, "fused_min_batch", b))) return rc;
        if ((rc = step_ms(b, &tf))) return rc;
        fused_wins.push_back(tf < ts);
        js += (i ? ", " : "") + std::string("\"") + std::to_string(b) + "\": [" + fmt("%.4f", ts) + ", " +
              fmt("%.4f", tf) + "]";
    }
    for (int i = (int)sizes.size() - 1; i >= 0 && fused_wins[i]; --i) fused_from = sizes[i];
    js += "}, \"fused_min_batch\": " + std::to_string(fused_from);
    if ((rc = kspir_context_set(c, "fused_min_batch", fused_from))) return rc;
    if (max_batch >= fused_from) {
        // the fused giant-step kernel: one workgroup per CU with rows loaded
        // through registers or prefetched by LDS-DMA, or two per CU
        double g1 = 0, gd = 0, g2 = 0;
        if ((rc = kspir_context_set(c, "giant_wgs_per_cu", 1))) return rc;
        if ((rc = kspir_context_set(c, "giant_dma", 0))) return rc;
        if ((rc = step_ms(max_batch, &g1))) return rc;
        if ((rc = kspir_context_set(c, "giant_dma", 1))) return rc;
        if ((rc = step_ms(max_batch, &gd))) return rc;
        if ((rc = kspir_context_set(c, "giant_dma", 0))) return rc;
        if ((rc = kspir_context_set(c, "giant_wgs_per_cu", 2))) return rc;
        if ((rc = step_ms(max_batch, &g2))) return rc;
        const int gw = g2 < std::min(g1, gd) ? 2 : 1, dma = gw == 1 && gd < g1 ? 1 : 0;
        js += ", \"giant_variant_ms\": {\"1\": " + fmt("%.4f", g1) + ", \"1_dma\": " + fmt("%.4f", gd) + ", \"2\": " +
              fmt("%.4f", g2) + "}, \"giant_wgs_per_cu\": " + std::to_string(gw) + ", \"giant_dma\": " +
              std::to_string(dma);
        if ((rc = kspir_context_set(c, "giant_wgs_per_cu", gw))) return rc;
        if ((rc = kspir_context_set(c, "giant_dma", dma))) return rc;
        double t1 = 0, t0 = 0;
        if ((rc = kspir_context_set(c, "fused_pack", 1))) return rc;
        if ((rc = step_ms(max_batch, &t1))) return rc;
        if ((rc = kspir_context_set(c, "fused_pack", 0))) return rc;
        if ((rc = step_ms(max_batch, &t0))) return rc;
        const int fp = t1 <= t0 ? 1 : 0;
        js += ", \"pack_ms\": [" + fmt("%.4f", t0) + ", " + fmt("%.4f", t1) + "], \"fused_pack\": " + std::to_string(fp);
        if ((rc = kspir_context_set(c, "fused_pack", fp))) return rc;
        if (fp) {
            // levels with few (pair, query) workgroups may do better as three kernels
            int64_t best_w = 0;
            double bw = t1;
            js += ", \"pack_fused_min_wgs_ms\": {\"0\": " + fmt("%.4f", t1);
            for (int64_t w : {64, 128, 256, 512}) {
                if ((rc = kspir_context_set(c, "pack_fused_min_wgs", w))) return rc;
                double t = 0;
                if ((rc = step_ms(max_batch, &t))) return rc;
                js += ", \"" + std::to_string(w) + "\": " + fmt("%.4f", t);
                if (t < bw) { bw = t; best_w = w; }
            }
            js += "}, \"pack_fused_min_wgs\": " + std::to_string(best_w);
            if ((rc = kspir_context_set(c, "pack_fused_min_wgs", best_w))) return rc;
        }
    }

    // 5. serving batch: highest queries/s within the latency budget
    int sb = 1;
    double best_qps = 0;
    js += ", \"step_ms\": {";
    for (size_t i = 0; i < sizes.size(); ++i) {
        const int b = sizes[i];
        double t = 0;
        if ((rc = step_ms(b, &t))) return rc;
        const double qps = b * 1000.0 / t;
        js += (i ? ", " : "") + std::string("\"") + std::to_string(b) + "\": " + fmt("%.4f", t);
        if ((latency_budget_ms <= 0 || t <= latency_budget_ms) && qps > best_qps) { best_qps = qps; sb = b; }
    }
    js += "}, \"serve_batch\": " + std::to_string(sb) + ", \"serve_qps\": " + fmt("%.1f", best_qps) + "}";
    if (serve_batch) *serve_batch = sb;
    if (report && report_len) {
        const size_t n = std::min(js.size(), report_len - 1);
        std::memcpy(report, js.data(), n);
        report[n] = 0;
    }
    return KSPIR_OK;
}
