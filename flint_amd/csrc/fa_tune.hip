// fa_tune.hip -- the plan space of integer products, plan features for the
// learned planner, and the exhaustive autotuner that times every plan.
//
// Candidate j < 64 of a product of an x bn limbs is (np, bits) with
// np = j / 8 + 1 and bits = bmax(np) - 8 (j mod 8), where bmax(np) is the
// largest chunk size np primes can carry exactly for this shape
// (fa_mpn_max_bits); np = 1 and sizes the constraints reject are invalid.
// The automatic plan (fa_mpn_choose_plan) is always candidate
// (np - 1) * 8 for its np.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fa_internal.h"
#include "../../include/flint_amd.h"

#define FA_NCAND 64

extern "C" int fa_mpn_candidates(int64_t an, int64_t bn, int32_t* out)
{
    if (an <= 0 || bn <= 0 || !out) return fa_set_error(FA_E_ARG, "fa_mpn_candidates: bad arguments");
    for (int j = 0; j < FA_NCAND; j++) {
        const int np = j / 8 + 1;
        int bits = 0;
        if (np >= 2) {
            const int bmax = fa_mpn_max_bits((uint64_t)an, (uint64_t)bn, np);
            bits = bmax ? bmax - 8 * (j % 8) : 0;
            if (bits > 0) {
                fa_mpn_plan_t P;
                // silent validity check: the same constraints as fa_mpn_fixed_plan
                fa_crt_consts C;
                fa_crt_consts_init(&C, np, 0);
                if (9 * bits < 64 * C.nl + 64) bits = 0;
                else if (fa_mpn_fixed_plan((uint64_t)an, (uint64_t)bn, np, bits, &P) != FA_OK) bits = 0;
            } else {
                bits = 0;
            }
        }
        out[2 * j] = bits ? np : 0;
        out[2 * j + 1] = bits;
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

// times every candidate of an x bn on the device alone: each candidate runs
// once outside any capture (twiddle tables and scratch reach the largest
// size any candidate needs), then K back-to-back products (K = 2^21 /
// (an + bn), between 1 and 64) of each are captured into a hipGraph of its
// own and replayed once untimed.  The timed replays are interleaved: replay
// r of every candidate runs before replay r + 1 of any, so that a slow
// stretch of the device (clocks, another process) lands on one sample of
// many candidates rather than on every sample of one.  ms[j] = the median
// replay time per product (HIP events around each replay), -1 for invalid
// candidates.  The first call of a process first keeps the device busy for
// about 50 ms so that its clocks have risen before any timing.  Host launch
// overhead is outside the measurement.
extern "C" int fa_autotune_mpn(int64_t an, int64_t bn, int reps, float* ms)
{
    if (an <= 0 || bn <= 0 || reps <= 0 || !ms)
        return fa_set_error(FA_E_ARG, "fa_autotune_mpn: bad arguments");
    int32_t cand[2 * FA_NCAND];
    int rc = fa_mpn_candidates(an, bn, cand);
    if (rc) return rc;
    fa_thread_state* st = fa_state();
    if (!st) return FA_E_NODEV;
    const size_t words = 2 * (size_t)(an + bn);
    if (words > tune_words) {
        if (tune_buf) (void)hipFree(tune_buf);
        tune_buf = nullptr;
        tune_words = 0;
        if (hipMalloc((void**)&tune_buf, words * 8) != hipSuccess)
            return fa_set_error(FA_E_NOMEM, "fa_autotune_mpn: %zu words", words);
        tune_words = words;
        (void)hipMemsetAsync(tune_buf, 0xA5, words * 8, st->stream);
        (void)hipStreamSynchronize(st->stream);
    }
    uint64_t* a = tune_buf;
    uint64_t* b = a + an;
    uint64_t* z = b + bn;
    int64_t K = (INT64_C(1) << 21) / (an + bn);
    if (K < 1) K = 1;
    if (K > 64) K = 64;
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
        return fa_set_error(FA_E_HIP, "fa_autotune_mpn: stream");
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return fa_set_error(FA_E_HIP, "fa_autotune_mpn: events");
    }
    static thread_local bool warmed = false;
    if (!warmed) {
        // clocks up: repeated products of the first valid candidate
        for (int j = 0; j < FA_NCAND; j++) {
            if (!cand[2 * j]) continue;
            float t = 0;
            (void)hipEventRecord(e0, cs);
            for (int r = 0; r < 1000 && t < 50.f && !rc; r++) {
                rc = fa_mpn_mul_device_plan(z, a, an, b, bn, cand[2 * j], cand[2 * j + 1], cs);
                (void)hipEventRecord(e1, cs);
                if (hipEventSynchronize(e1) != hipSuccess) break;
                (void)hipEventElapsedTime(&t, e0, e1);
            }
            break;
        }
        warmed = true;
    }
    // every candidate once outside the capture
    for (int j = 0; j < FA_NCAND && !rc; j++) {
        ms[j] = -1.f;
        if (!cand[2 * j]) continue;
        rc = fa_mpn_mul_device_plan(z, a, an, b, bn, cand[2 * j], cand[2 * j + 1], cs);
        if (!rc && hipStreamSynchronize(cs) != hipSuccess)
            rc = fa_set_error(FA_E_HIP, "tune warm-up");
    }
    std::vector<hipGraph_t> g((size_t)FA_NCAND, nullptr);
    std::vector<hipGraphExec_t> ex((size_t)FA_NCAND, nullptr);
    for (int j = 0; j < FA_NCAND && !rc; j++) {
        if (!cand[2 * j]) continue;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = fa_set_error(FA_E_HIP, "tune capture");
            break;
        }
        for (int64_t k = 0; k < K && !rc; k++)
            rc = fa_mpn_mul_device_plan(z, a, an, b, bn, cand[2 * j], cand[2 * j + 1], cs);
        const hipError_t ce = hipStreamEndCapture(cs, &g[(size_t)j]);
        if (!rc && (ce != hipSuccess || !g[(size_t)j] ||
                    hipGraphInstantiate(&ex[(size_t)j], g[(size_t)j], nullptr, nullptr, 0) !=
                        hipSuccess))
            rc = fa_set_error(FA_E_HIP, "tune graph");
        if (!rc && (hipGraphLaunch(ex[(size_t)j], cs) != hipSuccess ||
                    hipStreamSynchronize(cs) != hipSuccess))
            rc = fa_set_error(FA_E_HIP, "tune launch");
    }
    std::vector<float> times((size_t)FA_NCAND * (size_t)reps, 0.f);
    for (int r = 0; r < reps && !rc; r++) {
        for (int j = 0; j < FA_NCAND && !rc; j++) {
            if (!ex[(size_t)j]) continue;
            (void)hipEventRecord(e0, cs);
            if (hipGraphLaunch(ex[(size_t)j], cs) != hipSuccess) {
                rc = fa_set_error(FA_E_HIP, "tune launch");
                break;
            }
            (void)hipEventRecord(e1, cs);
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = fa_set_error(FA_E_HIP, "tune sync");
                break;
            }
            float t = 0;
            (void)hipEventElapsedTime(&t, e0, e1);
            times[(size_t)j * (size_t)reps + (size_t)r] = t;
        }
    }
    for (int j = 0; j < FA_NCAND; j++) {
        if (!rc && ex[(size_t)j]) {
            float* tj = &times[(size_t)j * (size_t)reps];
            std::sort(tj, tj + reps);
            const float med = (reps & 1) ? tj[reps / 2] : 0.5f * (tj[reps / 2 - 1] + tj[reps / 2]);
            ms[j] = med / (float)K;
        }
        if (ex[(size_t)j]) (void)hipGraphExecDestroy(ex[(size_t)j]);
        if (g[(size_t)j]) (void)hipGraphDestroy(g[(size_t)j]);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(cs);
    (void)hipStreamDestroy(cs);
    return rc;
}
