// fa_tune.hip -- the plan space of integer products, plan features for the
// learned planner, and the exhaustive autotuner that times every plan.
//
// Candidate j < 56 of a product of an x bn limbs is (np, bits, trunc) with
//     np = j / 8 + 2,  bits = bmax(np) - 8 ((j / 2) mod 4),  trunc = j mod 2,
// where bmax(np) is the largest chunk size np primes carry exactly for this
// shape (fa_mpn_max_bits) and trunc = 1 runs the transform truncated at the
// rows of the product (only where that applies: 2^13..2^23 points and a
// product ending before the last row).  Candidates 56..63 and sizes the
// constraints reject are invalid.  (Smaller chunks never won on the stored
// timings: the best plan of every one of 2,200 shapes had bits >= bmax - 24.)
// The automatic plan of a batch (fa_mpn_choose_plan, truncated where
// truncation applies) is candidate (np - 2) * 8 + trunc.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fa_internal.h"
#include "../../include/flint_amd.h"

#define FA_NCAND 64
#define FA_NVALID 56

// truncation applies to a plan: two-level transforms below 2^24 points
// whose product ends before the last row
static int trunc_applies(const fa_mpn_plan_t& P)
{
    if (P.depth <= FA_ROW_LG || P.depth >= 2 * FA_ROW_LG) return 0;
    const uint64_t rows = (P.zlen + (1u << FA_ROW_LG) - 1) >> FA_ROW_LG;
    return rows < ((uint64_t)1 << (P.depth - FA_ROW_LG));
}

extern "C" int fa_mpn_candidates(int64_t an, int64_t bn, int32_t* out)
{
    if (an <= 0 || bn <= 0 || !out) return fa_set_error(FA_E_ARG, "fa_mpn_candidates: bad arguments");
    for (int j = 0; j < FA_NCAND; j++) {
        out[3 * j] = out[3 * j + 1] = out[3 * j + 2] = 0;
        if (j >= FA_NVALID) continue;
        const int np = j / 8 + 2, off = (j / 2) % 4, tr = j % 2;
        const int bmax = fa_mpn_max_bits((uint64_t)an, (uint64_t)bn, np);
        const int bits = bmax ? bmax - 8 * off : 0;
        if (bits <= 0) continue;
        // silent validity check: the same constraints as fa_mpn_fixed_plan
        fa_crt_consts C;
        fa_crt_consts_init(&C, np, 0);
        if (9 * bits < 64 * C.nl + 64) continue;
        fa_mpn_plan_t P;
        if (fa_mpn_fixed_plan((uint64_t)an, (uint64_t)bn, np, bits, &P) != FA_OK) continue;
        if (tr && !trunc_applies(P)) continue;
        out[3 * j] = np;
        out[3 * j + 1] = bits;
        out[3 * j + 2] = tr;
    }
    return FA_OK;
}

// 16 features of a product shape for the planner
extern "C" int fa_plan_features(int64_t an, int64_t bn, float* f)
{
    if (an <= 0 || bn <= 0 || !f) return fa_set_error(FA_E_ARG, "fa_plan_features: bad arguments");
    const double la = log2((double)an), lb = log2((double)bn);
    fa_mpn_plan_t P;
    memset(&P, 0, sizeof(P));
    if (fa_mpn_choose_plan((uint64_t)an, (uint64_t)bn, &P) != FA_OK) memset(&P, 0, sizeof(P));
    f[0] = (float)(la / 24);
    f[1] = (float)(lb / 24);
    f[2] = (float)(fmin(la, lb) / 24);
    f[3] = (float)(fmax(la, lb) / 24);
    f[4] = (float)(log2((double)(an + bn)) / 25);
    f[5] = (float)((la + lb) / 48);
    f[6] = (float)(P.np / 8.0);
    f[7] = (float)(P.bits / 200.0);
    f[8] = (float)(P.depth / 24.0);
    f[9] = P.depth ? (float)((double)P.zlen / (double)(UINT64_C(1) << P.depth)) : 0.f;
    for (int np = 3; np <= 8; np++)
        f[7 + np] = (float)(fa_mpn_max_bits((uint64_t)an, (uint64_t)bn, np) / 200.0);
    return FA_OK;
}

// device buffers reused by the tuner (per host thread)
static thread_local uint64_t* tune_buf = nullptr;
static thread_local size_t tune_words = 0;

// times every candidate of an x bn as batched products: K copies of the
// product (K = 2^21 / (an + bn), between 1 and 64; shared random operands,
// one output each) form a device batch with the candidate's plan
// (fa_request np, bits, trunc), recorded into a hipGraph, replayed once
// untimed.  The operands are SplitMix64 limbs written afresh on every call
// (the carry stage's work depends on the data; stale buffer contents would
// leak one shape's timing into the next).  The timed replays are
// interleaved: replay r of every candidate runs before replay r + 1 of any,
// so that a slow stretch of the device (clocks, another process) lands on
// one sample of many candidates rather than on every sample of one.  ms[j] =
// the median replay time per product (HIP events around each replay), -1 for
// invalid candidates.  Candidates are timed in groups whose batches fit
// 16 GB of device workspace.  The first call of a process first keeps the
// device busy for about 50 ms so that its clocks have risen.  Host launch
// overhead is outside the measurement.
extern "C" int fa_autotune_mpn(int64_t an, int64_t bn, int reps, float* ms)
{
    if (an <= 0 || bn <= 0 || reps <= 0 || !ms)
        return fa_set_error(FA_E_ARG, "fa_autotune_mpn: bad arguments");
    int32_t cand[3 * FA_NCAND];
    int rc = fa_mpn_candidates(an, bn, cand);
    if (rc) return rc;
    for (int j = 0; j < FA_NCAND; j++) ms[j] = -1.f;
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    int64_t K = (INT64_C(1) << 21) / (an + bn);
    if (K < 1) K = 1;
    if (K > 64) K = 64;
    const size_t words = (size_t)(an + bn) * (size_t)(K + 1);
    if (words > tune_words) {
        if (tune_buf) (void)hipFree(tune_buf);
        tune_buf = nullptr;
        tune_words = 0;
        if (hipMalloc((void**)&tune_buf, words * 8) != hipSuccess)
            return fa_set_error(FA_E_NOMEM, "fa_autotune_mpn: %zu words", words);
        tune_words = words;
    }
    uint64_t* a = tune_buf;
    uint64_t* b = a + an;
    uint64_t* z = b + bn;
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
        return fa_set_error(FA_E_HIP, "fa_autotune_mpn: stream");
    fa_fill_spec F[2];
    memset(F, 0, sizeof(F));
    F[0].dst = a; F[0].count = (uint64_t)an; F[0].kind = FA_REQ_MPN;
    F[1].dst = b; F[1].count = (uint64_t)bn; F[1].kind = FA_REQ_MPN;
    F[0].key = ((uint64_t)an << 32) ^ (uint64_t)bn;
    F[1].key = F[0].key ^ UINT64_C(0x5DEECE66D);
    rc = fa_stream_fill_device(F, 2, cs);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fa_set_error(FA_E_HIP, "fa_autotune_mpn: events");
    static thread_local bool warmed = false;
    if (!rc && !warmed) {
        // clocks up: repeated products of the automatic plan
        float t = 0;
        (void)hipEventRecord(e0, cs);
        for (int r = 0; r < 1000 && t < 50.f && !rc; r++) {
            rc = fa_mpn_mul_device(z, a, an, b, bn, cs);
            (void)hipEventRecord(e1, cs);
            if (hipEventSynchronize(e1) != hipSuccess) break;
            (void)hipEventElapsedTime(&t, e0, e1);
        }
        warmed = true;
    }
    // one request array per candidate, batches created group by group
    std::vector<std::vector<fa_request>> R((size_t)FA_NCAND);
    std::vector<fa_batch> B((size_t)FA_NCAND, nullptr);
    std::vector<float> times((size_t)FA_NCAND * (size_t)reps, 0.f);
    int j0 = 0;
    while (!rc && j0 < FA_NCAND) {
        // a group: consecutive candidates up to 16 GB of batch workspace
        size_t ws = 0;
        int j1 = j0;
        for (; j1 < FA_NCAND; j1++) {
            if (!cand[3 * j1]) continue;
            fa_mpn_plan_t P;
            if (fa_mpn_fixed_plan((uint64_t)an, (uint64_t)bn, cand[3 * j1], cand[3 * j1 + 1], &P))
                continue;
            const size_t w = (size_t)K * (16 * (size_t)P.np * fa_res_stride(P.depth) * 8 +
                                          16 * (size_t)(an + bn) + 4096);
            if (ws && ws + w > (UINT64_C(16) << 30)) break;
            ws += w;
        }
        if (j1 == j0) j1 = j0 + 1;
        for (int j = j0; j < j1 && !rc; j++) {
            if (!cand[3 * j]) continue;
            auto& q = R[(size_t)j];
            q.assign((size_t)K, fa_request());
            for (int64_t k = 0; k < K; k++) {
                fa_request& r = q[(size_t)k];
                memset(&r, 0, sizeof(r));
                r.kind = FA_REQ_MPN;
                r.a = a;
                r.b = b;
                r.an = an;
                r.bn = bn;
                r.z = z + (size_t)k * (size_t)(an + bn);
                r.np = cand[3 * j];
                r.bits = cand[3 * j + 1];
                r.trunc = cand[3 * j + 2] ? FA_TRUNC_ON : FA_TRUNC_OFF;
            }
            rc = fa_batch_create(&B[(size_t)j], q.data(), K, FA_BATCH_DEVICE | FA_BATCH_GRAPH);
            for (int64_t k = 0; !rc && k < K; k++)
                if (q[(size_t)k].status != 1) rc = fa_set_error(FA_E_ARG, "tune: candidate %d", j);
            if (!rc) rc = fa_batch_launch(B[(size_t)j], cs);
            if (!rc && hipStreamSynchronize(cs) != hipSuccess)
                rc = fa_set_error(FA_E_HIP, "tune warm-up");
        }
        for (int r = 0; r < reps && !rc; r++) {
            for (int j = j0; j < j1 && !rc; j++) {
                if (!B[(size_t)j]) continue;
                (void)hipEventRecord(e0, cs);
                rc = fa_batch_launch(B[(size_t)j], cs);
                (void)hipEventRecord(e1, cs);
                if (!rc && hipEventSynchronize(e1) != hipSuccess)
                    rc = fa_set_error(FA_E_HIP, "tune sync");
                float t = 0;
                (void)hipEventElapsedTime(&t, e0, e1);
                times[(size_t)j * (size_t)reps + (size_t)r] = t;
            }
        }
        for (int j = j0; j < j1; j++) {
            if (!rc && B[(size_t)j]) {
                float* tj = &times[(size_t)j * (size_t)reps];
                std::sort(tj, tj + reps);
                const float med = (reps & 1) ? tj[reps / 2]
                                             : 0.5f * (tj[reps / 2 - 1] + tj[reps / 2]);
                ms[j] = med / (float)K;
            }
            if (B[(size_t)j]) fa_batch_destroy(B[(size_t)j]);
            B[(size_t)j] = nullptr;
        }
        j0 = j1;
    }
    for (auto& x : B)
        if (x) fa_batch_destroy(x);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(cs);
    (void)hipStreamDestroy(cs);
    return rc;
}
